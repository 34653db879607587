// The gate kinds of the fusion queue (host planner and device tile code).
#pragma once

namespace qa {
namespace fuse {

// F_DENSEN: an uncontrolled dense matrix on 3 or 4 qubits (a 4-qubit
// superoperator of a density matrix among them), applied to the tile in LDS
enum Kind : int { F_DENSE1 = 1, F_DENSE2, F_DIAG1, F_PARITY, F_PAULI, F_SWAP, F_DEPOL1, F_DEPH1, F_DENSE1V, F_DENSEN };

}  // namespace fuse
}  // namespace qa
