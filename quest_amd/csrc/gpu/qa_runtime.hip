// Device binding, memory, copies and timing of the kernel layer.
// Replaces quest/src/gpu/gpu_config.cpp (device query/bind :124-353,
// alloc :399-433, copies :449-606, cache :632-687) for one MI355X.
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "qa_common.hpp"
#include "qa_fused.hpp"

namespace {

thread_local char g_err[512] = "";
bool g_init = false;
int g_device = -1;
hipStream_t g_stream = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
double* g_devScratch = nullptr;   // reduction partials
size_t g_devScratchLen = 0;
double* g_hostScratch = nullptr;  // pinned landing area for reduction results
size_t g_hostScratchLen = 0;
double* g_devStage = nullptr;     // staging for strided density-matrix reads
size_t g_devStageBytes = 0;

// amplitude staging for precision-converting host transfers
constexpr size_t STAGE_AMPS = 1u << 22;  // 4 Mi amplitudes = 64 MiB at FP64

template <typename S, typename D>
__global__ void k_convert(D* __restrict__ dst, const S* __restrict__ src, uint64_t numReals) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < numReals; i += stride) dst[i] = (D)src[i];
}

// gather a rows x cols block of a column-major density matrix into a
// row-major FP64 buffer
template <typename T>
__global__ void k_gatherBlock(double* __restrict__ out, const T* __restrict__ amps, uint64_t dim,
                              uint64_t r0, uint64_t c0, uint64_t nr, uint64_t nc) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = nr * nc;
    if (k >= total) return;
    uint64_t r = k / nc, c = k % nc;
    uint64_t i = (r0 + r) + (c0 + c) * dim;
    out[2 * k] = (double)amps[2 * i];
    out[2 * k + 1] = (double)amps[2 * i + 1];
}

}  // namespace

namespace qa {

void setError(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
hipStream_t stream() { return g_stream; }
bool isInit() { return g_init; }
int device() { return g_device; }

std::atomic<uint64_t> g_epoch{1};
void bumpEpoch() { g_epoch.fetch_add(1, std::memory_order_relaxed); }

double* devScratch(size_t n) {
    if (n > g_devScratchLen) {
        if (g_devScratch) (void)hipFree(g_devScratch);
        size_t len = n < (1u << 16) ? (1u << 16) : n;
        if (hipMalloc(&g_devScratch, len * sizeof(double)) != hipSuccess) {
            g_devScratch = nullptr;
            g_devScratchLen = 0;
            return nullptr;
        }
        g_devScratchLen = len;
    }
    return g_devScratch;
}

// a reduction that needed an unusually large scratch (an outcome table of
// many qubits) gives it back rather than hold it until the library ends
void trimDevScratch() {
    constexpr size_t KEEP = 1u << 22;  // 32 MiB
    if (g_devScratchLen > KEEP) {
        (void)hipStreamSynchronize(stream());
        (void)hipFree(g_devScratch);
        g_devScratch = nullptr;
        g_devScratchLen = 0;
    }
}

double* hostScratch(size_t n) {
    if (n > g_hostScratchLen) {
        if (g_hostScratch) (void)hipHostFree(g_hostScratch);
        size_t len = n < 4096 ? 4096 : n;
        if (hipHostMalloc(&g_hostScratch, len * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            g_hostScratch = nullptr;
            g_hostScratchLen = 0;
            return nullptr;
        }
        g_hostScratchLen = len;
    }
    return g_hostScratch;
}

}  // namespace qa

extern "C" {

const char* qa_lastError(void) { return g_err; }

int qa_isGpuAvailable(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0;
}

uint64_t qa_stateEpoch(void) { return qa::g_epoch.load(std::memory_order_relaxed); }

int qa_init(int device) {
    qa::bumpEpoch();
    if (g_init) {
        if (device == g_device || device < 0) return QA_OK;
        qa::setError("qa_init: already bound to device %d", g_device);
        return QA_ERR_ARG;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        qa::setError("qa_init: no HIP device is visible");
        return QA_ERR_NO_DEVICE;
    }
    if (device < 0) device = 0;
    QA_ARG_CHECK(device < n, "qa_init: device %d out of range (%d visible)", device, n);
    QA_HIP(hipSetDevice(device));
    // default priority stream: every entry point enqueues here
    QA_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    QA_HIP(hipEventCreate(&g_ev0));
    QA_HIP(hipEventCreate(&g_ev1));
    if (!qa::devScratch(1u << 16) || !qa::hostScratch(4096)) {
        qa::setError("qa_init: scratch allocation failed");
        return QA_ERR_OOM;
    }
    {
        const int rc = qa::initNoAllocScratch(device);
        if (rc != QA_OK) return rc;
    }
    g_device = device;
    g_init = true;
    return QA_OK;
}

int qa_finalize(void) {
    qa::bumpEpoch();
    if (!g_init) return QA_OK;
    QA_FLUSH_ALL();
    QA_HIP(hipStreamSynchronize(g_stream));
    qa::fuse::shutdown();
    if (g_devScratch) QA_HIP(hipFree(g_devScratch));
    if (g_hostScratch) QA_HIP(hipHostFree(g_hostScratch));
    if (g_devStage) QA_HIP(hipFree(g_devStage));
    g_devStage = nullptr;
    g_devStageBytes = 0;
    g_devScratch = nullptr;
    g_hostScratch = nullptr;
    g_devScratchLen = g_hostScratchLen = 0;
    QA_HIP(hipEventDestroy(g_ev0));
    QA_HIP(hipEventDestroy(g_ev1));
    QA_HIP(hipStreamDestroy(g_stream));
    g_stream = nullptr;
    g_init = false;
    g_device = -1;
    return QA_OK;
}

int qa_getMemInfo(size_t* freeBytes, size_t* totalBytes) {
    QA_REQUIRE_INIT();
    QA_HIP(hipMemGetInfo(freeBytes, totalBytes));
    return QA_OK;
}

int qa_getDeviceInfo(int* numCUs, char* name, int nameLen) {
    QA_REQUIRE_INIT();
    hipDeviceProp_t p;
    QA_HIP(hipGetDeviceProperties(&p, g_device));
    if (numCUs) *numCUs = p.multiProcessorCount;
    if (name && nameLen > 0) {
        snprintf(name, (size_t)nameLen, "%s (%s)", p.name, p.gcnArchName);
    }
    return QA_OK;
}

int qa_malloc(void** ptr, size_t numBytes) {
    QA_REQUIRE_INIT();
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, numBytes);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        *ptr = nullptr;
        qa::setError("qa_malloc: %zu bytes do not fit in device memory", numBytes);
        return QA_ERR_OOM;
    }
    QA_HIP(e);
    return QA_OK;
}

int qa_free(void* ptr) {
    QA_REQUIRE_INIT();
    if (!ptr) return QA_OK;
    qa::fuse::drop(ptr);  // gates still queued on a freed state are discarded
    QA_HIP(hipStreamSynchronize(g_stream));
    QA_HIP(hipFree(ptr));
    return QA_OK;
}

int qa_memcpyHtoD(void* dst, const void* src, size_t n) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, g_stream));
    QA_HIP(hipStreamSynchronize(g_stream));
    return QA_OK;
}

int qa_memcpyDtoH(void* dst, const void* src, size_t n) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, g_stream));
    QA_HIP(hipStreamSynchronize(g_stream));
    return QA_OK;
}

int qa_memcpyDtoD(void* dst, const void* src, size_t n) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, g_stream));
    return QA_OK;
}

int qa_sync(void) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipStreamSynchronize(g_stream));
    return QA_OK;
}

void* qa_getStream(void) { return (void*)g_stream; }

int qa_getAmps(double* out, const void* amps, int prec, int64_t start, int64_t num) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_ARG_CHECK(start >= 0 && num >= 0, "qa_getAmps: negative range");
    if (num == 0) return QA_OK;
    if (prec == QA_FP64) {
        QA_HIP(hipMemcpyAsync(out, (const double*)amps + 2 * start, (size_t)num * 16,
                              hipMemcpyDeviceToHost, g_stream));
        QA_HIP(hipStreamSynchronize(g_stream));
        return QA_OK;
    }
    QA_ARG_CHECK(prec == QA_FP32, "qa_getAmps: bad precision %d", prec);
    std::vector<float> stage;
    for (int64_t done = 0; done < num;) {
        int64_t n = num - done;
        if (n > (int64_t)STAGE_AMPS) n = (int64_t)STAGE_AMPS;
        stage.resize((size_t)(2 * n));
        QA_HIP(hipMemcpyAsync(stage.data(), (const float*)amps + 2 * (start + done), (size_t)n * 8,
                              hipMemcpyDeviceToHost, g_stream));
        QA_HIP(hipStreamSynchronize(g_stream));
        for (int64_t k = 0; k < 2 * n; k++) out[2 * done + k] = (double)stage[(size_t)k];
        done += n;
    }
    return QA_OK;
}

int qa_setAmps(void* amps, int prec, int64_t start, const double* in, int64_t num) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_ARG_CHECK(start >= 0 && num >= 0, "qa_setAmps: negative range");
    if (num == 0) return QA_OK;
    if (prec == QA_FP64) {
        QA_HIP(hipMemcpyAsync((double*)amps + 2 * start, in, (size_t)num * 16, hipMemcpyHostToDevice,
                              g_stream));
        QA_HIP(hipStreamSynchronize(g_stream));
        return QA_OK;
    }
    QA_ARG_CHECK(prec == QA_FP32, "qa_setAmps: bad precision %d", prec);
    std::vector<float> stage;
    for (int64_t done = 0; done < num;) {
        int64_t n = num - done;
        if (n > (int64_t)STAGE_AMPS) n = (int64_t)STAGE_AMPS;
        stage.resize((size_t)(2 * n));
        for (int64_t k = 0; k < 2 * n; k++) stage[(size_t)k] = (float)in[2 * done + k];
        QA_HIP(hipMemcpyAsync((float*)amps + 2 * (start + done), stage.data(), (size_t)n * 8,
                              hipMemcpyHostToDevice, g_stream));
        QA_HIP(hipStreamSynchronize(g_stream));
        done += n;
    }
    return QA_OK;
}

int qa_getDensityAmps(double* out, const void* amps, int prec, int numQubits, int64_t r0, int64_t c0,
                      int64_t nr, int64_t nc) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    int64_t dim = 1ll << numQubits;
    QA_ARG_CHECK(r0 >= 0 && c0 >= 0 && nr >= 0 && nc >= 0 && r0 + nr <= dim && c0 + nc <= dim,
                 "qa_getDensityAmps: block out of range");
    uint64_t total = (uint64_t)nr * (uint64_t)nc;
    if (!total) return QA_OK;
    const uint64_t chunkRows = (STAGE_AMPS / (uint64_t)nc) ? (STAGE_AMPS / (uint64_t)nc) : 1;
    size_t bufAmps = (size_t)(chunkRows * (uint64_t)nc);
    if (g_devStageBytes < bufAmps * 16) {
        // grown once, kept until qa_finalize(): no allocation on later calls
        if (g_devStage) QA_HIP(hipFree(g_devStage));
        g_devStage = nullptr;
        g_devStageBytes = 0;
        QA_HIP(hipMalloc(&g_devStage, bufAmps * 16));
        g_devStageBytes = bufAmps * 16;
    }
    double* dbuf = g_devStage;
    int rc = QA_OK;
    for (uint64_t r = 0; r < (uint64_t)nr && rc == QA_OK; r += chunkRows) {
        uint64_t rows = ((uint64_t)nr - r < chunkRows) ? ((uint64_t)nr - r) : chunkRows;
        uint64_t cnt = rows * (uint64_t)nc;
        unsigned g = (unsigned)((cnt + 255) / 256);
        if (prec == QA_FP64)
            k_gatherBlock<double><<<g, 256, 0, g_stream>>>(dbuf, (const double*)amps, (uint64_t)dim,
                                                            (uint64_t)r0 + r, (uint64_t)c0, rows,
                                                            (uint64_t)nc);
        else
            k_gatherBlock<float><<<g, 256, 0, g_stream>>>(dbuf, (const float*)amps, (uint64_t)dim,
                                                           (uint64_t)r0 + r, (uint64_t)c0, rows,
                                                           (uint64_t)nc);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out + 2 * r * (uint64_t)nc, dbuf, cnt * 16, hipMemcpyDeviceToHost,
                           g_stream) != hipSuccess ||
            hipStreamSynchronize(g_stream) != hipSuccess) {
            qa::setError("qa_getDensityAmps: gather failed");
            rc = QA_ERR_HIP;
        }
    }
    (void)hipStreamSynchronize(g_stream);
    return rc;
}

int qa_copyAmps(void* dst, int dstPrec, const void* src, int srcPrec, int64_t num) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    if (num <= 0) return QA_OK;
    uint64_t reals = 2ull * (uint64_t)num;
    if (dstPrec == srcPrec) {
        size_t bytes = (size_t)reals * (dstPrec == QA_FP64 ? 8 : 4);
        QA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
        return QA_OK;
    }
    unsigned g = (unsigned)((reals + 255) / 256);
    if (g > 65536) g = 65536;
    if (dstPrec == QA_FP64)
        k_convert<float, double><<<g, 256, 0, g_stream>>>((double*)dst, (const float*)src, reals);
    else
        k_convert<double, float><<<g, 256, 0, g_stream>>>((float*)dst, (const double*)src, reals);
    QA_LAUNCH_CHECK();
    return QA_OK;
}

int qa_timerStart(void) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipEventRecord(g_ev0, g_stream));
    return QA_OK;
}

int qa_eventCreate(void** ev) {
    QA_REQUIRE_INIT();
    hipEvent_t e;
    QA_HIP(hipEventCreate(&e));
    *ev = (void*)e;
    return QA_OK;
}

int qa_eventRecord(void* ev) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipEventRecord((hipEvent_t)ev, g_stream));
    return QA_OK;
}

int qa_eventElapsedMs(void* start, void* stop, double* ms) {
    QA_REQUIRE_INIT();
    QA_HIP(hipEventSynchronize((hipEvent_t)stop));
    float f = 0;
    QA_HIP(hipEventElapsedTime(&f, (hipEvent_t)start, (hipEvent_t)stop));
    *ms = (double)f;
    return QA_OK;
}

int qa_eventDestroy(void* ev) {
    QA_REQUIRE_INIT();
    QA_HIP(hipEventDestroy((hipEvent_t)ev));
    return QA_OK;
}

int qa_timerStop(double* ms) {
    QA_REQUIRE_INIT();
    QA_FLUSH_ALL();
    QA_HIP(hipEventRecord(g_ev1, g_stream));
    QA_HIP(hipEventSynchronize(g_ev1));
    float f = 0;
    QA_HIP(hipEventElapsedTime(&f, g_ev0, g_ev1));
    *ms = (double)f;
    return QA_OK;
}

}  // extern "C"
