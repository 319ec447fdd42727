// C ABI of the gfx950 frame loop (declarations and contracts in include/azu_hip.h).
// Owns the handles, validates arguments, turns HIP failures into status codes and
// thread-local messages, and optionally brackets every launch with HIP events for the
// per-kernel timing report.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/azu_hip.h"
#include "../../include/tracking/TrackingPolicy.h"
#include "common.hpp"

namespace azu {
hipError_t launchIntegrate(int format, bool exact, void* vox, const VolParams& p, Intr K, int W, int H,
                           const float* depth, float min_d, float max_d, const PoseArg& pa, const FixedDiv& dtr,
                           bool recip, void* bounds, unsigned long long* counters, unsigned char* bricks,
                           unsigned char* classes, hipStream_t st);
size_t depthBoundsBytes(int W, int H);
size_t exactScratchBytes(int res);
hipError_t launchFillEmpty(int format, void* vox, size_t n, hipStream_t st);
hipError_t launchBrickMap(int format, const void* vox, int res, unsigned char* bricks, bool scan, hipStream_t st);
hipError_t launchRaycast(int format, bool exact, const void* vox, const VolParams& p, const FixedDiv& dvs,
                         const unsigned char* bricks, unsigned* dsafe, int* ok, const PoseArg& pa, Intr K, int W,
                         int H, float* verts, float* norms, float* t_hit, hipStream_t st);
hipError_t launchSkipInit(unsigned* dsafe, size_t cells, hipStream_t st);
hipError_t launchLive(int src_kind, const void* src, int W, int H, Intr K, float min_d, float max_d, float* const d[3],
                      float* const v[3], float* const n[3], hipStream_t st);
size_t icpStateBytes();
int icpBlocksFor(int W, int H);
size_t icpPartialsBytes(int maxBlocks);
hipError_t launchIcpInit(void* st, const PoseArg& est, const PoseArg& ref, hipStream_t s);
hipError_t launchPredict(const float* cur, float* last, float* pred, hipStream_t s);
hipError_t launchWritePose(void* st, const PoseArg& p, hipStream_t s);
hipError_t launchIcpClearTicks(void* st, hipStream_t s);
int depthLossBlocks(int W, int H);
struct McResult {
    unsigned long long ntri, nvert;
    bool truncated;
    float* pos;
    float* nrm;
    uint32_t* idx;
};
hipError_t launchPointCloud(int format, const void* vox, const VolParams& p, float* d_out,
                            unsigned long long capacity, unsigned long long* total, hipStream_t st);
hipError_t launchMarchingCubes(int format, const void* vox, const VolParams& p, unsigned long long max_tris,
                               McResult* res, hipStream_t st);
hipError_t launchDepthLoss(int format, const void* vox, const VolParams& p, const PoseArg& pa, Intr K, int W, int H,
                           const float* d_obs, float step, int backward, float* grad_tsdf, float* depth_out,
                           double* partials, double* out8, const unsigned char* bricks, unsigned* dsafe, int* ok,
                           hipStream_t st);
void icpTicksRead(const void* hostCopy, unsigned long long* ticks);
hipError_t launchIcpIter(const float* lv, const float* ln, int W, int H, const float* mv, const float* mn, int MW,
                         int MH, Intr Km, float angle_cos, float dist_sq, int level, int iter, int prev_level,
                         int prev_iter, int prev_W, int prev_H, int k, void* st, double* partials, hipStream_t s);
void icpStateRead(const void* hostCopy, float* pose, float* information, float* error, float* final_step,
                  int* inliers, int* converged, int* counters, int* iterations);
hipError_t launchDownsample(int W, int H, float* const d[3], float* const v[3], float* const n[3], hipStream_t st);
hipError_t launchVolConvert(int format, bool to_float, void* vox, float* tsdf, float* weight, size_t n,
                            hipStream_t st);
hipError_t launchCountObserved(int format, const void* vox, size_t n, unsigned long long* out, hipStream_t st);
hipError_t launchCondStage(int stage, const uint16_t* in, const float* luma, uint16_t* out, float* state, int W,
                           int H, float lo, float hi, hipStream_t st);
hipError_t launchCondZero(float* p, size_t n, hipStream_t st);
}  // namespace azu

using namespace azu;

namespace {

// Integration counters: 256 slot pairs of u64, summed when read (integrate.hip); the
// allocation holds one more u64 after them for azu_vol_usage.
constexpr size_t kCounterBytes = 256 * 2 * sizeof(unsigned long long);
constexpr size_t kCounterAlloc = kCounterBytes + 64;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int hipFail(hipError_t e, const char* what) {
    return fail(AZU_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define AZU_HIP_TRY(expr, what)                          \
    do {                                                 \
        const hipError_t e_ = (expr);                    \
        if (e_ != hipSuccess) return hipFail(e_, what);  \
    } while (0)

// ---- per-kernel timing ----------------------------------------------------------
struct ProfEntry {
    long long launches = 0;
    double total_ms = 0.0;
};
struct ProfPending {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::map<std::string, ProfEntry> g_prof;
std::vector<ProfPending> g_pending;

struct ProfScope {
    hipStream_t s;
    const char* name;
    hipEvent_t a{}, b{};
    bool on;
    ProfScope(const char* n, hipStream_t st) : s(st), name(n) {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        on = g_prof_on;
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, s);
        }
    }
    ~ProfScope() {
        if (!on) return;
        (void)hipEventRecord(b, s);
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_pending.push_back(ProfPending{name, a, b});
    }
};

void profCollect() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& p : g_pending) {
        (void)hipEventSynchronize(p.b);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ProfEntry& e = g_prof[p.name];
            e.launches += 1;
            e.total_ms += ms;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_pending.clear();
}

bool validParams(const azu_tsdf_params* p, std::string* why) {
    if (!p) { *why = "null params"; return false; }
    if (p->resolution < 2 || p->resolution > 2048) { *why = "resolution must be in [2, 2048]"; return false; }
    if (!(p->voxel_size > 0.0f) || !std::isfinite(p->voxel_size)) { *why = "voxel_size must be > 0"; return false; }
    if (!(p->truncation > 0.0f) || !std::isfinite(p->truncation)) { *why = "truncation must be > 0"; return false; }
    if (!(p->max_weight >= 1.0f) || !(p->max_weight <= 65535.0f) || p->max_weight != std::floor(p->max_weight)) {
        *why = "max_weight must be an integer in [1, 65535]";
        return false;
    }
    for (float o : p->origin)
        if (!std::isfinite(o)) { *why = "origin must be finite"; return false; }
    return true;
}

VolParams toVol(const azu_tsdf_params& p) {
    VolParams v;
    v.res = p.resolution;
    v.vs = p.voxel_size;
    v.trunc = p.truncation;
    v.max_w = p.max_weight;
    v.origin[0] = p.origin[0];
    v.origin[1] = p.origin[1];
    v.origin[2] = p.origin[2];
    v.min_depth = p.min_depth;
    v.max_depth = p.max_depth;
    return v;
}

bool sameParams(const azu_tsdf_params& a, const azu_tsdf_params& b) {
    return a.resolution == b.resolution && a.voxel_size == b.voxel_size && a.truncation == b.truncation &&
           a.max_weight == b.max_weight && a.min_depth == b.min_depth && a.max_depth == b.max_depth &&
           a.origin[0] == b.origin[0] && a.origin[1] == b.origin[1] && a.origin[2] == b.origin[2];
}

PoseArg poseByValue(const float* m) {
    PoseArg pa;
    for (int i = 0; i < 16; ++i) pa.m[i] = m[i];
    pa.dptr = nullptr;
    return pa;
}
PoseArg poseByPtr(const float* d) {
    PoseArg pa;
    for (int i = 0; i < 16; ++i) pa.m[i] = 0.0f;
    pa.dptr = d;
    return pa;
}
bool finiteK(const float* K) {
    return K && std::isfinite(K[0]) && std::isfinite(K[1]) && std::isfinite(K[2]) && std::isfinite(K[3]) &&
           K[0] != 0.0f && K[1] != 0.0f;
}

// scaleIntrinsics (include/sensor/CameraIntrinsics.h:33-40)
Intr scaleIntrinsics(Intr k, int rw, int rh, int w, int h) {
    if (w == rw && h == rh) return k;
    const float sx = (float)w / (float)rw, sy = (float)h / (float)rh;
    return Intr{k.fx * sx, k.fy * sy, (k.cx + 0.5f) * sx - 0.5f, (k.cy + 0.5f) * sy - 0.5f};
}

}  // namespace

// Device scratch of azu_integrate_noalloc: a file-scope array every GPU holds its own
// copy of until the process ends. Host code never passes the array's name to a kernel
// (on the host it is only a stand-in for the symbol); it passes the device address that
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with that
// device current.
int scratchAddress(void** out) {
    int dev = 0;
    AZU_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(AZU_ERR_STATE, "device id out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hipFail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return AZU_OK;
}
}  // namespace

struct azu_volume {
    azu_tsdf_params params;
    int format;
    int mode;
    void* vox = nullptr;
    size_t bytes = 0;
    unsigned long long* counters = nullptr;  // device: 256 slot pairs {updated, newly observed}
    double* adj_partials = nullptr;          // depth-loss block partials
    size_t adj_bytes = 0;
    double* adj_out = nullptr;               // depth-loss result [8]
    int frames = 0;
    FixedDiv dvs{1.0f, 1.0f, 0}, dtr{1.0f, 1.0f, 0};  // division by voxel size / truncation
    bool recip = false;                                // verified Newton-Raphson 1 / x
    void* bounds = nullptr;                            // depth bounds of the last frame size
    size_t bounds_bytes = 0;
    unsigned char* classes = nullptr;                  // exact integration's tile-class map + row carries
    size_t classes_bytes = 0;
    // Brick map for the raycast's empty-space skipping, kept by reset / upload /
    // integrate. Once the raw storage has been handed out, other writers may exist and
    // the map is rebuilt from the voxels before every raycast.
    unsigned char* bricks = nullptr;
    bool external = false;
    bool bricks_stale = false;  // voxels written by azu_vol_write_voxels since the last map rebuild
    bool skip = true;  // empty-space skipping in the raycast
    unsigned* dsafe = nullptr;  // skip bounds, two words per 2 x 2 pixel cell, then one flag word
    size_t dsafe_cells = 0;
    bool exact() const { return mode != AZU_MODE_FAST; }
};

// Depth conditioner: two ping-pong raw frames for the stage chain and the EMA history.
struct azu_conditioner {
    int W = 0, H = 0;
    int minimal = 0;
    uint16_t* buf[2] = {nullptr, nullptr};
    float* ema = nullptr;
};

struct azu_tracker {
    int W, H;
    Intr K;
    azu_icp_params params;
    int lw[3], lh[3];
    float* d[3] = {nullptr, nullptr, nullptr};
    float* v[3] = {nullptr, nullptr, nullptr};
    float* n[3] = {nullptr, nullptr, nullptr};
    void* state = nullptr;
    double* partials = nullptr;
    void* host_state = nullptr;  // pinned
    float* motion = nullptr;     // device: previous pose [16], predicted pose [16]
    bool tracked = false;
};

namespace {
azu_icp_params sanitize(const azu_icp_params& p) {
    azu_icp_params o = p;
    // sanitizeParams (src/tracking/ICPTracker.cpp:24-34)
    if (!std::isfinite(o.angle_threshold)) o.angle_threshold = 30.0f;
    else if (o.angle_threshold < 0.0f) o.angle_threshold = 0.0f;
    else if (o.angle_threshold > 85.0f) o.angle_threshold = 85.0f;
    return o;
}

int volAlloc(azu_volume* v) {
    const size_t n = (size_t)v->params.resolution * v->params.resolution * v->params.resolution;
    const size_t wb = v->format == AZU_FMT_F32 ? 8 : 4;
    v->bytes = n * wb;
    const hipError_t e = hipMalloc(&v->vox, v->bytes);
    if (e != hipSuccess) {
        v->vox = nullptr;
        return fail(AZU_ERR_ALLOC, std::string("volume allocation of ") + std::to_string(v->bytes) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    const size_t nb = (size_t)brickCount(v->params.resolution);
    if (hipMalloc(&v->bricks, nb * nb * nb) != hipSuccess) {
        v->bricks = nullptr;
        (void)hipFree(v->vox);
        v->vox = nullptr;
        return fail(AZU_ERR_ALLOC, "brick map allocation failed");
    }
    return AZU_OK;
}
}  // namespace

extern "C" {

const char* azu_last_error(void) { return g_err.c_str(); }
int azu_abi_version(void) { return AZU_HIP_ABI_VERSION; }

int azu_device_info(int* device, char* name, size_t name_len, size_t* hbm_bytes) {
    int dev = 0;
    AZU_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    AZU_HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (device) *device = dev;
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return AZU_OK;
}

int azu_malloc(void** p, size_t bytes) {
    if (!p) return fail(AZU_ERR_ARG, "null output pointer");
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(AZU_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    return AZU_OK;
}
int azu_free(void* p) {
    AZU_HIP_TRY(hipFree(p), "hipFree");
    return AZU_OK;
}
int azu_host_alloc(void** p, size_t bytes) {
    if (!p) return fail(AZU_ERR_ARG, "null output pointer");
    const hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(AZU_ERR_ALLOC, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    return AZU_OK;
}
int azu_host_free(void* p) {
    AZU_HIP_TRY(hipHostFree(p), "hipHostFree");
    return AZU_OK;
}
int azu_memcpy_h2d(void* d, const void* s, size_t b, azu_stream st) {
    AZU_HIP_TRY(hipMemcpyAsync(d, s, b, hipMemcpyHostToDevice, (hipStream_t)st), "hipMemcpyAsync H2D");
    return AZU_OK;
}
int azu_memcpy_d2h(void* d, const void* s, size_t b, azu_stream st) {
    AZU_HIP_TRY(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToHost, (hipStream_t)st), "hipMemcpyAsync D2H");
    return AZU_OK;
}
int azu_memcpy_d2d(void* d, const void* s, size_t b, azu_stream st) {
    AZU_HIP_TRY(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, (hipStream_t)st), "hipMemcpyAsync D2D");
    return AZU_OK;
}
int azu_stream_create(azu_stream* s) {
    if (!s) return fail(AZU_ERR_ARG, "null output pointer");
    hipStream_t h;
    AZU_HIP_TRY(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreate");
    *s = h;
    return AZU_OK;
}
int azu_stream_destroy(azu_stream s) {
    AZU_HIP_TRY(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy");
    return AZU_OK;
}
int azu_stream_sync(azu_stream s) {
    AZU_HIP_TRY(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
    return AZU_OK;
}

// ---- volume ----------------------------------------------------------------------
int azu_vol_create(const azu_tsdf_params* p, int format, azu_volume** out) {
    if (!out) return fail(AZU_ERR_ARG, "null output pointer");
    *out = nullptr;
    std::string why;
    if (!validParams(p, &why)) return fail(AZU_ERR_ARG, why);
    if (format != AZU_FMT_F32 && format != AZU_FMT_F16_U16) return fail(AZU_ERR_ARG, "unknown volume format");
    auto* v = new azu_volume();
    v->params = *p;
    v->format = format;
    v->mode = AZU_MODE_AUTO;
    int rc = volAlloc(v);
    if (rc != AZU_OK) {
        delete v;
        return rc;
    }
    if (hipMalloc(&v->counters, kCounterAlloc) != hipSuccess) {
        (void)hipFree(v->vox);
        (void)hipFree(v->bricks);
        delete v;
        return fail(AZU_ERR_ALLOC, "counter allocation failed");
    }
    rc = azu_vol_reset(v, nullptr);
    if (rc == AZU_OK) {
        const hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = hipFail(e, "volume reset");
    }
    v->dvs = makeFixedDiv(v->params.voxel_size);
    v->dtr = makeFixedDiv(v->params.truncation);
    v->recip = recipVerified();
    if (rc != AZU_OK) {
        azu_vol_destroy(v);
        return rc;
    }
    *out = v;
    return AZU_OK;
}

int azu_vol_destroy(azu_volume* v) {
    if (!v) return AZU_OK;
    if (v->vox) (void)hipFree(v->vox);
    if (v->bricks) (void)hipFree(v->bricks);
    if (v->dsafe) (void)hipFree(v->dsafe);
    if (v->counters) (void)hipFree(v->counters);
    if (v->adj_partials) (void)hipFree(v->adj_partials);
    if (v->adj_out) (void)hipFree(v->adj_out);
    if (v->bounds) (void)hipFree(v->bounds);
    if (v->classes) (void)hipFree(v->classes);
    delete v;
    return AZU_OK;
}

int azu_vol_reset(azu_volume* v, azu_stream s) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    const size_t n = (size_t)v->params.resolution * v->params.resolution * v->params.resolution;
    {
        ProfScope ps("fill_empty", (hipStream_t)s);
        AZU_HIP_TRY(launchFillEmpty(v->format, v->vox, n, (hipStream_t)s), "fill_empty launch");
        AZU_HIP_TRY(launchBrickMap(v->format, v->vox, v->params.resolution, v->bricks, false, (hipStream_t)s),
                    "brick map clear");
    }
    AZU_HIP_TRY(hipMemsetAsync(v->counters, 0, kCounterBytes, (hipStream_t)s), "counter reset");
    v->frames = 0;
    v->bricks_stale = false;
    return AZU_OK;
}

int azu_vol_set_params(azu_volume* v, const azu_tsdf_params* p) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    std::string why;
    if (!validParams(p, &why)) return fail(AZU_ERR_ARG, why);
    if (sameParams(*p, v->params)) return AZU_OK;
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before setParams");
    if (p->resolution != v->params.resolution) {
        // The new storage is allocated before the old one is released: if it fails, the
        // volume keeps its storage, parameters and contents, and stays usable.
        azu_volume fresh;
        fresh.params = *p;
        fresh.format = v->format;
        const int rc = volAlloc(&fresh);
        if (rc != AZU_OK) return rc;
        (void)hipFree(v->vox);
        (void)hipFree(v->bricks);
        v->vox = fresh.vox;
        v->bricks = fresh.bricks;
        v->bytes = fresh.bytes;
    }
    v->params = *p;
    const int rc = azu_vol_reset(v, nullptr);
    if (rc != AZU_OK) return rc;
    AZU_HIP_TRY(hipStreamSynchronize(nullptr), "volume reset");
    v->dvs = makeFixedDiv(v->params.voxel_size);
    v->dtr = makeFixedDiv(v->params.truncation);
    return AZU_OK;
}

int azu_vol_division_check(const azu_volume* v, int* voxel_size_fast, int* truncation_fast, int* reciprocal_fast) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    if (voxel_size_fast) *voxel_size_fast = v->dvs.fast;
    if (truncation_fast) *truncation_fast = v->dtr.fast;
    if (reciprocal_fast) *reciprocal_fast = v->recip ? 1 : 0;
    return AZU_OK;
}

int azu_vol_get_params(const azu_volume* v, azu_tsdf_params* p, int* format) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    if (p) *p = v->params;
    if (format) *format = v->format;
    return AZU_OK;
}

int azu_vol_set_mode(azu_volume* v, int mode) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    if (mode != AZU_MODE_AUTO && mode != AZU_MODE_EXACT && mode != AZU_MODE_FAST)
        return fail(AZU_ERR_ARG, "unknown kernel mode");
    if (mode == AZU_MODE_FAST && v->params.resolution % 4 != 0)
        return fail(AZU_ERR_ARG, "the fast kernels need a resolution divisible by 4");
    v->mode = mode;
    return AZU_OK;
}

int azu_vol_set_skip(azu_volume* v, int on) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    v->skip = on != 0;
    return AZU_OK;
}

static int volTransfer(azu_volume* v, float* tsdf, float* weight, bool download) {
    if (!v || !tsdf || !weight) return fail(AZU_ERR_ARG, "null argument");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    // Work on this volume may be queued on any stream (non-blocking ones included), which
    // the null stream used below does not wait for.
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before volume transfer");
    const size_t n = (size_t)v->params.resolution * v->params.resolution * v->params.resolution;
    const size_t chunk = std::min<size_t>(n, (size_t)1 << 24);
    float *dt = nullptr, *dw = nullptr;
    if (hipMalloc(&dt, chunk * 4) != hipSuccess || hipMalloc(&dw, chunk * 4) != hipSuccess) {
        if (dt) (void)hipFree(dt);
        return fail(AZU_ERR_ALLOC, "staging allocation failed");
    }
    int rc = AZU_OK;
    const size_t wb = v->format == AZU_FMT_F32 ? 8 : 4;
    for (size_t off = 0; off < n && rc == AZU_OK; off += chunk) {
        const size_t m = std::min(chunk, n - off);
        char* base = static_cast<char*>(v->vox) + off * wb;
        hipError_t e = hipSuccess;
        if (download) {
            e = launchVolConvert(v->format, true, base, dt, dw, m, nullptr);
            if (e == hipSuccess) e = hipMemcpy(tsdf + off, dt, m * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(weight + off, dw, m * 4, hipMemcpyDeviceToHost);
        } else {
            e = hipMemcpy(dt, tsdf + off, m * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(dw, weight + off, m * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = launchVolConvert(v->format, false, base, dt, dw, m, nullptr);
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        }
        if (e != hipSuccess) rc = hipFail(e, download ? "volume download" : "volume upload");
    }
    if (!download && rc == AZU_OK) {
        hipError_t e = launchBrickMap(v->format, v->vox, v->params.resolution, v->bricks, true, nullptr);
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = hipFail(e, "brick map rebuild");
        else v->bricks_stale = false;
    }
    (void)hipFree(dt);
    (void)hipFree(dw);
    return rc;
}

int azu_vol_download(azu_volume* v, float* tsdf, float* weight) { return volTransfer(v, tsdf, weight, true); }
int azu_vol_upload(azu_volume* v, const float* tsdf, const float* weight) {
    return volTransfer(v, const_cast<float*>(tsdf), const_cast<float*>(weight), false);
}

namespace {
// Host form of the storage words (the device's StoreF32 / StoreF16 layout).
float wordTsdf(int format, const unsigned char* w) {
    if (format == AZU_FMT_F32) {
        float t;
        std::memcpy(&t, w, 4);
        return t;
    }
    uint16_t h;
    std::memcpy(&h, w, 2);
    return __half2float(__ushort_as_half(h));
}
float wordWeight(int format, const unsigned char* w) {
    if (format == AZU_FMT_F32) {
        float x;
        std::memcpy(&x, w + 4, 4);
        return x;
    }
    uint16_t h;
    std::memcpy(&h, w + 2, 2);
    return (float)h;
}
}  // namespace

int azu_vol_read_voxels(azu_volume* v, size_t first, size_t count, float* tsdf, float* weight) {
    if (!v || (count && (!tsdf || !weight))) return fail(AZU_ERR_ARG, "null argument");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    const size_t n = (size_t)v->params.resolution * v->params.resolution * v->params.resolution;
    if (first > n || count > n - first) return fail(AZU_ERR_ARG, "voxel range outside the volume");
    if (!count) return AZU_OK;
    const size_t wb = v->format == AZU_FMT_F32 ? 8 : 4;
    std::vector<unsigned char> h(count * wb);
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before voxel read");
    AZU_HIP_TRY(hipMemcpy(h.data(), static_cast<char*>(v->vox) + first * wb, count * wb, hipMemcpyDeviceToHost),
                "voxel read");
    for (size_t i = 0; i < count; ++i) {
        tsdf[i] = wordTsdf(v->format, &h[i * wb]);
        weight[i] = wordWeight(v->format, &h[i * wb]);
    }
    return AZU_OK;
}

int azu_vol_write_voxels(azu_volume* v, size_t first, size_t count, const float* tsdf, const float* weight) {
    if (!v || (count && (!tsdf || !weight))) return fail(AZU_ERR_ARG, "null argument");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    const size_t n = (size_t)v->params.resolution * v->params.resolution * v->params.resolution;
    if (first > n || count > n - first) return fail(AZU_ERR_ARG, "voxel range outside the volume");
    if (!count) return AZU_OK;
    const size_t wb = v->format == AZU_FMT_F32 ? 8 : 4;
    std::vector<unsigned char> h(count * wb);
    for (size_t i = 0; i < count; ++i) {
        unsigned char* w = &h[i * wb];
        if (v->format == AZU_FMT_F32) {
            std::memcpy(w, &tsdf[i], 4);
            std::memcpy(w + 4, &weight[i], 4);
        } else {
            const float wt = weight[i];
            if (!(wt >= 0.0f && wt <= 65535.0f) || wt != std::floor(wt))
                return fail(AZU_ERR_ARG, "an F16_U16 volume stores integer weights in [0, 65535]");
            const uint16_t ht = __half_as_ushort(__float2half_rn(tsdf[i]));
            const uint16_t hw = (uint16_t)wt;
            std::memcpy(w, &ht, 2);
            std::memcpy(w + 2, &hw, 2);
        }
    }
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before voxel write");
    AZU_HIP_TRY(hipMemcpy(static_cast<char*>(v->vox) + first * wb, h.data(), count * wb, hipMemcpyHostToDevice),
                "voxel write");
    v->bricks_stale = true;  // the next raycast rebuilds the brick map from the voxels
    return AZU_OK;
}

int azu_vol_usage(azu_volume* v, unsigned long long* observed) {
    if (!v || !observed) return fail(AZU_ERR_ARG, "null argument");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    const size_t n = (size_t)v->params.resolution * v->params.resolution * v->params.resolution;
    unsigned long long* slot = v->counters + kCounterBytes / sizeof(unsigned long long);
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before usage count");
    AZU_HIP_TRY(hipMemset(slot, 0, sizeof(unsigned long long)), "usage count clear");
    AZU_HIP_TRY(launchCountObserved(v->format, v->vox, n, slot, nullptr), "usage count launch");
    AZU_HIP_TRY(hipMemcpy(observed, slot, sizeof(unsigned long long), hipMemcpyDeviceToHost), "usage count read");
    return AZU_OK;
}

int azu_vol_storage(azu_volume* v, void** d_ptr, size_t* bytes) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    if (d_ptr) *d_ptr = v->vox;
    if (bytes) *bytes = v->bytes;
    v->external = true;
    return AZU_OK;
}

int azu_vol_counters(azu_volume* v, unsigned long long* updated, unsigned long long* newly, int reset) {
    if (!v) return fail(AZU_ERR_ARG, "null volume");
    std::vector<unsigned long long> h(kCounterBytes / sizeof(unsigned long long));
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise for counters");
    AZU_HIP_TRY(hipMemcpy(h.data(), v->counters, kCounterBytes, hipMemcpyDeviceToHost), "counter read");
    unsigned long long u = 0, nw = 0;
    for (size_t i = 0; i + 1 < h.size(); i += 2) {
        u += h[i];
        nw += h[i + 1];
    }
    if (updated) *updated = u;
    if (newly) *newly = nw;
    if (reset) AZU_HIP_TRY(hipMemset(v->counters, 0, kCounterBytes), "counter reset");
    return AZU_OK;
}

int azu_vol_integrated_frames(const azu_volume* v, int* frames) {
    if (!v || !frames) return fail(AZU_ERR_ARG, "null argument");
    *frames = v->frames;
    return AZU_OK;
}

// bounds_scratch: null to use (and grow) the volume's own depth-bounds buffer, else a
// caller-provided buffer of at least depthBoundsBytes(W, H).
static int integrateImpl(azu_volume* v, const float* d_depth, int W, int H, const float* K, const PoseArg& pa,
                         float min_d, float max_d, azu_stream s, void* bounds_scratch = nullptr) {
    if (!v || !d_depth) return fail(AZU_ERR_ARG, "null volume or depth");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    if (W <= 0 || H <= 0) return fail(AZU_ERR_ARG, "image size must be positive");
    if (!finiteK(K)) return fail(AZU_ERR_ARG, "intrinsics must be finite with non-zero focal lengths");
    const bool exact = v->exact();
    if (!exact && v->params.resolution % 4 != 0)
        return fail(AZU_ERR_ARG, "the fast kernels need a resolution divisible by 4");
    ProfScope ps(exact ? (v->format == AZU_FMT_F32 ? "integrate_exact_f32" : "integrate_exact_f16")
                       : (v->format == AZU_FMT_F32 ? "integrate_fast_f32" : "integrate_fast_f16"),
                 (hipStream_t)s);
    if (!bounds_scratch) {
        const size_t need = depthBoundsBytes(W, H);
        if (need > v->bounds_bytes) {
            // grows only when the frame size does; the old buffer may still be read by
            // work queued on another stream, so wait for the device first
            AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before depth bounds reallocation");
            if (v->bounds) (void)hipFree(v->bounds);
            v->bounds = nullptr;
            v->bounds_bytes = 0;
            if (hipMalloc(&v->bounds, need) != hipSuccess) return fail(AZU_ERR_ALLOC, "depth bounds allocation failed");
            v->bounds_bytes = need;
        }
        const size_t cneed = exactScratchBytes(v->params.resolution);
        if (exact && cneed > v->classes_bytes) {
            AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before tile-class map / row carry reallocation");
            if (v->classes) (void)hipFree(v->classes);
            v->classes = nullptr;
            v->classes_bytes = 0;
            if (hipMalloc((void**)&v->classes, cneed) != hipSuccess)
                return fail(AZU_ERR_ALLOC, "tile-class map / row carry allocation failed");
            v->classes_bytes = cneed;
        }
    }
    AZU_HIP_TRY(launchIntegrate(v->format, exact, v->vox, toVol(v->params), Intr{K[0], K[1], K[2], K[3]}, W, H,
                                d_depth, min_d, max_d, pa, v->dtr, v->recip,
                                bounds_scratch ? bounds_scratch : v->bounds, v->counters, v->bricks,
                                bounds_scratch ? nullptr : v->classes, (hipStream_t)s),
                "integrate launch");
    v->frames += 1;
    return AZU_OK;
}

int azu_integrate(azu_volume* v, const float* d_depth_m, int W, int H, const float K[4], const float pose[16],
                  float min_depth, float max_depth, azu_stream s) {
    if (!pose) return fail(AZU_ERR_ARG, "null pose");
    return integrateImpl(v, d_depth_m, W, H, K, poseByValue(pose), min_depth, max_depth, s);
}

int azu_integrate_dpose(azu_volume* v, const float* d_depth_m, int W, int H, const float K[4], const float* d_pose,
                        float min_depth, float max_depth, azu_stream s) {
    if (!d_pose) return fail(AZU_ERR_ARG, "null device pose");
    return integrateImpl(v, d_depth_m, W, H, K, poseByPtr(d_pose), min_depth, max_depth, s);
}

int azu_integrate_noalloc(azu_volume* v, const float* d_depth_m, int W, int H, const float K[4], const float pose[16],
                          float min_depth, float max_depth, azu_stream s) {
    if (!pose) return fail(AZU_ERR_ARG, "null pose");
    if (!v || !d_depth_m) return fail(AZU_ERR_ARG, "null volume or depth");
    if (W <= 0 || H <= 0) return fail(AZU_ERR_ARG, "image size must be positive");
    if ((long long)W * H > (1LL << 26)) return fail(AZU_ERR_ARG, "frame too large");
    const size_t need = depthBoundsBytes(W, H);
    if (need > kNoAllocScratchBytes)
        return fail(AZU_ERR_ARG, "azu_integrate_noalloc: the depth bounds of a " + std::to_string(W) + "x" +
                                     std::to_string(H) + " frame need " + std::to_string(need) +
                                     " bytes, more than the " + std::to_string(kNoAllocScratchBytes) +
                                     "-byte scratch");
    void* sp = nullptr;
    const int rc = scratchAddress(&sp);
    if (rc != AZU_OK) return rc;
    return integrateImpl(v, d_depth_m, W, H, K, poseByValue(pose), min_depth, max_depth, s, sp);
}

// The skip bounds' scratch for a W x H image (grown with the image) and a current brick
// map (rebuilt from the voxels once other writers may have changed them).
static int prepareSkip(azu_volume* v, int W, int H, hipStream_t st) {
    const size_t cells = (size_t)((W + 1) / 2) * (size_t)((H + 1) / 2);
    if (cells > v->dsafe_cells) {
        // grows only with the image; work queued on other streams may still use the old one
        AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before skip scratch reallocation");
        if (v->dsafe) (void)hipFree(v->dsafe);
        v->dsafe = nullptr;
        v->dsafe_cells = 0;
        if (hipMalloc(&v->dsafe, (2 * cells + 1) * sizeof(unsigned)) != hipSuccess)
            return fail(AZU_ERR_ALLOC, "raycast skip scratch allocation failed");
        AZU_HIP_TRY(launchSkipInit(v->dsafe, cells, st), "skip scratch init");
        v->dsafe_cells = cells;
    }
    if (v->skip && (v->external || v->bricks_stale)) {
        AZU_HIP_TRY(launchBrickMap(v->format, v->vox, v->params.resolution, v->bricks, true, st), "brick map rebuild");
        v->bricks_stale = false;
    }
    return AZU_OK;
}

static int raycastImpl(azu_volume* v, const PoseArg& pa, const float* K, int W, int H, float* dv, float* dn,
                       float* dt, azu_stream s) {
    if (!v || !dv || !dn) return fail(AZU_ERR_ARG, "null volume or output");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    if (W <= 0 || H <= 0) return fail(AZU_ERR_ARG, "image size must be positive");
    if (!finiteK(K)) return fail(AZU_ERR_ARG, "intrinsics must be finite with non-zero focal lengths");
    const bool exact = v->exact();
    ProfScope ps(exact ? (v->format == AZU_FMT_F32 ? "raycast_exact_f32" : "raycast_exact_f16")
                       : (v->format == AZU_FMT_F32 ? "raycast_fast_f32" : "raycast_fast_f16"),
                 (hipStream_t)s);
    const int rc = prepareSkip(v, W, H, (hipStream_t)s);
    if (rc != AZU_OK) return rc;
    AZU_HIP_TRY(launchRaycast(v->format, exact, v->vox, toVol(v->params), v->dvs, v->skip ? v->bricks : nullptr,
                              v->dsafe,
                              reinterpret_cast<int*>(v->dsafe + 2 * v->dsafe_cells), pa, Intr{K[0], K[1], K[2], K[3]}, W,
                              H, dv, dn, dt, (hipStream_t)s),
                "raycast launch");
    return AZU_OK;
}

int azu_raycast(azu_volume* v, const float pose[16], const float K[4], int W, int H, float* d_vertices,
                float* d_normals, azu_stream s) {
    if (!pose) return fail(AZU_ERR_ARG, "null pose");
    return raycastImpl(v, poseByValue(pose), K, W, H, d_vertices, d_normals, nullptr, s);
}

int azu_raycast_dpose(azu_volume* v, const float* d_pose, const float K[4], int W, int H, float* d_vertices,
                      float* d_normals, float* d_t_hit, azu_stream s) {
    if (!d_pose) return fail(AZU_ERR_ARG, "null device pose");
    return raycastImpl(v, poseByPtr(d_pose), K, W, H, d_vertices, d_normals, d_t_hit, s);
}

int azu_raycast_depth_loss(azu_volume* v, const float pose[16], const float K[4], int W, int H,
                           const float* d_depth_obs, float step_voxels, int backward, float* d_grad_tsdf,
                           float* d_depth_out, double out[8], azu_stream s) {
    if (!v || !pose || !out) return fail(AZU_ERR_ARG, "null volume, pose or output");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    if (W <= 0 || H <= 0) return fail(AZU_ERR_ARG, "image size must be positive");
    if (!finiteK(K)) return fail(AZU_ERR_ARG, "intrinsics must be finite with non-zero focal lengths");
    if (!(step_voxels >= 0.05f && step_voxels <= 8.0f)) return fail(AZU_ERR_ARG, "step_voxels must be in [0.05, 8]");
    const hipStream_t st = (hipStream_t)s;
    const size_t need = sizeof(double) * 8 * (size_t)depthLossBlocks(W, H);
    if (v->adj_bytes < need) {
        if (v->adj_partials) (void)hipFree(v->adj_partials);
        v->adj_partials = nullptr;
        v->adj_bytes = 0;
        if (hipMalloc((void**)&v->adj_partials, need) != hipSuccess) return fail(AZU_ERR_ALLOC, "adjoint scratch");
        v->adj_bytes = need;
    }
    if (!v->adj_out && hipMalloc((void**)&v->adj_out, 8 * sizeof(double)) != hipSuccess)
        return fail(AZU_ERR_ALLOC, "adjoint output");
    {
        ProfScope ps(backward ? "depth_loss_bwd" : "depth_loss_fwd", st);
        const int rc = prepareSkip(v, W, H, st);
        if (rc != AZU_OK) return rc;
        AZU_HIP_TRY(launchDepthLoss(v->format, v->vox, toVol(v->params), poseByValue(pose),
                                    Intr{K[0], K[1], K[2], K[3]}, W, H, d_depth_obs, step_voxels * v->params.voxel_size,
                                    backward, d_grad_tsdf, d_depth_out, v->adj_partials, v->adj_out,
                                    v->skip ? v->bricks : nullptr, v->dsafe,
                                    reinterpret_cast<int*>(v->dsafe + 2 * v->dsafe_cells), st),
                    "depth loss launch");
    }
    AZU_HIP_TRY(hipMemcpyAsync(out, v->adj_out, 8 * sizeof(double), hipMemcpyDeviceToHost, st), "loss read");
    AZU_HIP_TRY(hipStreamSynchronize(st), "loss sync");
    return AZU_OK;
}

// ---- marching cubes --------------------------------------------------------------
struct azu_mesh {
    McResult r;
};

int azu_vol_point_cloud(azu_volume* v, float* d_points, unsigned long long capacity, unsigned long long* count,
                        azu_stream s) {
    if (!v || !count) return fail(AZU_ERR_ARG, "null volume or count");
    if (capacity > 0 && !d_points) return fail(AZU_ERR_ARG, "null point buffer with a non-zero capacity");
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    if ((size_t)v->params.resolution * v->params.resolution * v->params.resolution >= ((size_t)1 << 31))
        return fail(AZU_ERR_ARG, "point cloud: resolution^3 must stay below 2^31");
    // The null stream does not wait for work queued on non-blocking streams.
    if (!s) AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before point cloud");
    ProfScope ps("point_cloud", (hipStream_t)s);
    AZU_HIP_TRY(launchPointCloud(v->format, v->vox, toVol(v->params), d_points, capacity, count, (hipStream_t)s),
                "point cloud");
    return AZU_OK;
}

int azu_mc_extract(azu_volume* v, unsigned long long max_triangles, azu_mesh** out, azu_stream s) {
    if (!v || !out) return fail(AZU_ERR_ARG, "null volume or output");
    *out = nullptr;
    if (!v->vox) return fail(AZU_ERR_STATE, "volume has no storage");
    // The null stream does not wait for work queued on non-blocking streams.
    if (!s) AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise before marching cubes");
    auto* m = new azu_mesh();
    hipError_t e;
    {
        ProfScope ps("marching_cubes", (hipStream_t)s);
        e = launchMarchingCubes(v->format, v->vox, toVol(v->params), max_triangles, &m->r, (hipStream_t)s);
    }
    if (e != hipSuccess) {
        delete m;
        return hipFail(e, "marching cubes");
    }
    *out = m;
    return AZU_OK;
}

int azu_mesh_info(const azu_mesh* m, unsigned long long* vertices, unsigned long long* triangles, int* truncated) {
    if (!m) return fail(AZU_ERR_ARG, "null mesh");
    if (vertices) *vertices = m->r.nvert;
    if (triangles) *triangles = m->r.ntri;
    if (truncated) *truncated = m->r.truncated ? 1 : 0;
    return AZU_OK;
}

int azu_mesh_download(const azu_mesh* m, float* positions, float* normals, uint32_t* indices) {
    if (!m) return fail(AZU_ERR_ARG, "null mesh");
    if (positions && m->r.nvert)
        AZU_HIP_TRY(hipMemcpy(positions, m->r.pos, m->r.nvert * 12, hipMemcpyDeviceToHost), "mesh download");
    if (normals && m->r.nvert)
        AZU_HIP_TRY(hipMemcpy(normals, m->r.nrm, m->r.nvert * 12, hipMemcpyDeviceToHost), "mesh download");
    if (indices && m->r.ntri)
        AZU_HIP_TRY(hipMemcpy(indices, m->r.idx, m->r.ntri * 12, hipMemcpyDeviceToHost), "mesh download");
    return AZU_OK;
}

int azu_mesh_device(const azu_mesh* m, const float** p, const float** n, const uint32_t** i) {
    if (!m) return fail(AZU_ERR_ARG, "null mesh");
    if (p) *p = m->r.pos;
    if (n) *n = m->r.nrm;
    if (i) *i = m->r.idx;
    return AZU_OK;
}

int azu_mesh_destroy(azu_mesh* m) {
    if (!m) return AZU_OK;
    if (m->r.pos) (void)hipFree(m->r.pos);
    if (m->r.nrm) (void)hipFree(m->r.nrm);
    if (m->r.idx) (void)hipFree(m->r.idx);
    delete m;
    return AZU_OK;
}

// ---- tracker ---------------------------------------------------------------------
int azu_tracker_destroy(azu_tracker* t) {
    if (!t) return AZU_OK;
    for (int l = 0; l < 3; ++l) {
        if (t->d[l]) (void)hipFree(t->d[l]);
        if (t->v[l]) (void)hipFree(t->v[l]);
        if (t->n[l]) (void)hipFree(t->n[l]);
    }
    if (t->state) (void)hipFree(t->state);
    if (t->partials) (void)hipFree(t->partials);
    if (t->host_state) (void)hipHostFree(t->host_state);
    if (t->motion) (void)hipFree(t->motion);
    delete t;
    return AZU_OK;
}

int azu_tracker_create(int W, int H, const float K[4], const azu_icp_params* p, azu_tracker** out) {
    if (!out) return fail(AZU_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (W < 4 || H < 4 || W > 8192 || H > 8192) return fail(AZU_ERR_ARG, "live frame size must be in [4, 8192]");
    if (!finiteK(K)) return fail(AZU_ERR_ARG, "intrinsics must be finite with non-zero focal lengths");
    if (!p) return fail(AZU_ERR_ARG, "null ICP params");
    auto* t = new azu_tracker();
    t->W = W;
    t->H = H;
    t->K = Intr{K[0], K[1], K[2], K[3]};
    t->params = sanitize(*p);
    t->lw[0] = W;
    t->lh[0] = H;
    for (int l = 1; l < 3; ++l) {
        t->lw[l] = t->lw[l - 1] / 2;
        t->lh[l] = t->lh[l - 1] / 2;
    }
    bool ok = true;
    for (int l = 0; l < 3 && ok; ++l) {
        const size_t np = (size_t)t->lw[l] * t->lh[l];
        ok = hipMalloc(&t->d[l], np * 4) == hipSuccess && hipMalloc(&t->v[l], np * 12) == hipSuccess &&
             hipMalloc(&t->n[l], np * 12) == hipSuccess;
    }
    ok = ok && hipMalloc(&t->state, icpStateBytes()) == hipSuccess;
    ok = ok && hipMalloc((void**)&t->partials, icpPartialsBytes(1024)) == hipSuccess;
    ok = ok && hipHostMalloc(&t->host_state, icpStateBytes(), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMalloc((void**)&t->motion, 32 * sizeof(float)) == hipSuccess;
    if (!ok) {
        azu_tracker_destroy(t);
        return fail(AZU_ERR_ALLOC, "tracker allocation failed");
    }
    if (hipMemset(t->state, 0, icpStateBytes()) != hipSuccess ||
        hipMemset(t->partials, 0, icpPartialsBytes(1024)) != hipSuccess) {
        azu_tracker_destroy(t);
        return fail(AZU_ERR_HIP, "tracker initialisation failed");
    }
    *out = t;
    return AZU_OK;
}

int azu_tracker_set_params(azu_tracker* t, const azu_icp_params* p) {
    if (!t || !p) return fail(AZU_ERR_ARG, "null argument");
    t->params = sanitize(*p);
    return AZU_OK;
}

int azu_tracker_get_params(const azu_tracker* t, azu_icp_params* out) {
    if (!t || !out) return fail(AZU_ERR_ARG, "null argument");
    *out = t->params;
    return AZU_OK;
}

int azu_icp_params_sanitize(const azu_icp_params* in, azu_icp_params* out) {
    if (!in || !out) return fail(AZU_ERR_ARG, "null argument");
    *out = sanitize(*in);
    return AZU_OK;
}

int azu_icp_result_default(azu_icp_result* out) {
    if (!out) return fail(AZU_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    for (int i = 0; i < 4; ++i) out->pose[i * 5] = 1.0f;
    out->final_step = std::numeric_limits<float>::infinity();
    return AZU_OK;
}

static int prepareImpl(azu_tracker* t, const void* src, int src_kind, azu_stream s) {
    if (!t || !src) return fail(AZU_ERR_ARG, "null tracker or depth");
    ProfScope ps("live_pyramid", (hipStream_t)s);
    AZU_HIP_TRY(launchLive(src_kind, src, t->W, t->H, t->K, t->params.min_depth, t->params.max_depth, t->d, t->v, t->n,
                           (hipStream_t)s),
                "live pyramid launch");
    return AZU_OK;
}
int azu_tracker_prepare_mm(azu_tracker* t, const uint16_t* d, azu_stream s) { return prepareImpl(t, d, 1, s); }
int azu_tracker_prepare_m(azu_tracker* t, const float* d, azu_stream s) { return prepareImpl(t, d, 0, s); }
int azu_tracker_prepare_raw(azu_tracker* t, const uint16_t* d, azu_stream s) { return prepareImpl(t, d, 2, s); }

int azu_tracker_load_level(azu_tracker* t, int level, const float* d_depth, const float* d_vertices,
                           const float* d_normals, azu_stream s) {
    if (!t || level < 0 || level > 2 || !d_vertices || !d_normals) return fail(AZU_ERR_ARG, "bad tracker, level or buffer");
    const size_t np = (size_t)t->lw[level] * t->lh[level];
    const hipStream_t st = (hipStream_t)s;
    if (d_depth) AZU_HIP_TRY(hipMemcpyAsync(t->d[level], d_depth, np * 4, hipMemcpyDeviceToDevice, st), "level copy");
    AZU_HIP_TRY(hipMemcpyAsync(t->v[level], d_vertices, np * 12, hipMemcpyDeviceToDevice, st), "level copy");
    AZU_HIP_TRY(hipMemcpyAsync(t->n[level], d_normals, np * 12, hipMemcpyDeviceToDevice, st), "level copy");
    return AZU_OK;
}

int azu_tracker_downsample(azu_tracker* t, azu_stream s) {
    if (!t) return fail(AZU_ERR_ARG, "null tracker");
    ProfScope ps("live_pyramid", (hipStream_t)s);
    AZU_HIP_TRY(launchDownsample(t->W, t->H, t->d, t->v, t->n, (hipStream_t)s), "downsample launch");
    return AZU_OK;
}

int azu_tracker_level(azu_tracker* t, int level, const float** d, const float** v, const float** n, int* W, int* H) {
    if (!t || level < 0 || level > 2) return fail(AZU_ERR_ARG, "bad tracker or level");
    if (d) *d = t->d[level];
    if (v) *v = t->v[level];
    if (n) *n = t->n[level];
    if (W) *W = t->lw[level];
    if (H) *H = t->lh[level];
    return AZU_OK;
}

static int trackImpl(azu_tracker* t, const float* mv, const float* mn, int MW, int MH, const PoseArg& est,
                     const PoseArg& ref, azu_stream s) {
    if (!t || !mv || !mn) return fail(AZU_ERR_ARG, "null tracker or model");
    if (MW <= 0 || MH <= 0) return fail(AZU_ERR_ARG, "model size must be positive");
    const hipStream_t st = (hipStream_t)s;
    const Intr Km = scaleIntrinsics(t->K, t->W, t->H, MW, MH);
    // cos(angle * pi / 180) evaluated in double and rounded to float (ICPTracker.cpp:160)
    const float angle_cos = (float)std::cos(t->params.angle_threshold * M_PI / 180.0f);
    const float dist_sq = t->params.dist_threshold * t->params.dist_threshold;
    {
        ProfScope ps("icp_init", st);
        AZU_HIP_TRY(launchIcpInit(t->state, est, ref, st), "ICP init launch");
    }
    // Launch k solves iteration k - 1 and runs the pass of iteration k; a last launch
    // solves the final iteration.
    int k = 0, pl = -1, pi = 0;
    for (int level = 2; level >= 0; --level) {
        const int iters = t->params.max_iterations[level];
        for (int it = 0; it < iters; ++it, ++k) {
            ProfScope ps("icp_iter", st);
            AZU_HIP_TRY(launchIcpIter(t->v[level], t->n[level], t->lw[level], t->lh[level], mv, mn, MW, MH, Km,
                                      angle_cos, dist_sq, level, it, pl, pi, pl >= 0 ? t->lw[pl] : 0,
                                      pl >= 0 ? t->lh[pl] : 0, k, t->state, t->partials, st),
                        "ICP iteration launch");
            pl = level;
            pi = it;
        }
    }
    if (pl >= 0) {
        ProfScope ps("icp_iter", st);
        AZU_HIP_TRY(launchIcpIter(nullptr, nullptr, 0, 0, mv, mn, MW, MH, Km, angle_cos, dist_sq, -1, 0, pl, pi,
                                  t->lw[pl], t->lh[pl], k, t->state, t->partials, st),
                    "ICP final solve launch");
    }
    t->tracked = true;
    return AZU_OK;
}

int azu_tracker_track(azu_tracker* t, const float* mv, const float* mn, int MW, int MH, const float est[16],
                      const float ref[16], azu_stream s) {
    if (!est || !ref) return fail(AZU_ERR_ARG, "null pose");
    return trackImpl(t, mv, mn, MW, MH, poseByValue(est), poseByValue(ref), s);
}

int azu_tracker_track_dpose(azu_tracker* t, const float* mv, const float* mn, int MW, int MH, const float* d_est,
                            const float* d_ref, azu_stream s) {
    if (!d_est || !d_ref) return fail(AZU_ERR_ARG, "null device pose");
    return trackImpl(t, mv, mn, MW, MH, poseByPtr(d_est), poseByPtr(d_ref), s);
}

int azu_tracker_set_pose(azu_tracker* t, const float pose[16], azu_stream s) {
    if (!t || !pose) return fail(AZU_ERR_ARG, "null argument");
    ProfScope ps("icp_init", (hipStream_t)s);
    AZU_HIP_TRY(launchIcpInit(t->state, poseByValue(pose), poseByValue(pose), (hipStream_t)s), "pose set launch");
    // the motion model's previous pose starts equal to the current one (zero motion)
    AZU_HIP_TRY(hipMemcpyAsync(t->motion, t->state, 64, hipMemcpyDeviceToDevice, (hipStream_t)s), "pose set copy");
    return AZU_OK;
}

int azu_tracker_predict(azu_tracker* t, const float** d_predicted, azu_stream s) {
    if (!t) return fail(AZU_ERR_ARG, "null tracker");
    ProfScope ps("predict_pose", (hipStream_t)s);
    AZU_HIP_TRY(launchPredict(static_cast<const float*>(t->state), t->motion, t->motion + 16, (hipStream_t)s),
                "predict launch");
    if (d_predicted) *d_predicted = t->motion + 16;
    return AZU_OK;
}

int azu_tracker_set_tracked_pose(azu_tracker* t, const float pose[16], azu_stream s) {
    if (!t || !pose) return fail(AZU_ERR_ARG, "null argument");
    AZU_HIP_TRY(launchWritePose(t->state, poseByValue(pose), (hipStream_t)s), "pose write launch");
    return AZU_OK;
}

int azu_tracking_policy_defaults(azu_tracking_policy* p) {
    if (!p) return fail(AZU_ERR_ARG, "null output pointer");
    const kfusion::tracking::TrackingPolicy d;
    p->min_inliers = d.min_inliers;
    p->min_good_inliers = d.min_good_inliers;
    p->min_fit_ratio = d.min_fit_ratio;
    p->max_rms_m = d.max_rms_m;
    p->max_frame_translation = d.max_frame_translation;
    p->max_frame_rotation = d.max_frame_rotation;
    p->failures_before_lost = d.failures_before_lost;
    return AZU_OK;
}

int azu_track_grade(const azu_icp_result* r, const float prev_pose[16], const azu_tracking_policy* p,
                    float degeneracy_rel, azu_track_decision* out) {
    namespace T = kfusion::tracking;
    if (!r || !prev_pose || !out) return fail(AZU_ERR_ARG, "null argument");
    if (!(degeneracy_rel >= 0.0f) || !std::isfinite(degeneracy_rel))
        return fail(AZU_ERR_ARG, "degeneracy_rel must be finite and >= 0");
    T::TrackingPolicy pol;
    if (p) {
        pol.min_inliers = p->min_inliers;
        pol.min_good_inliers = p->min_good_inliers;
        pol.min_fit_ratio = p->min_fit_ratio;
        pol.max_rms_m = p->max_rms_m;
        pol.max_frame_translation = p->max_frame_translation;
        pol.max_frame_rotation = p->max_frame_rotation;
        pol.failures_before_lost = p->failures_before_lost;
    }
    T::ICPResult res;
    std::memcpy(res.pose.data(), r->pose, sizeof(r->pose));
    std::memcpy(res.information.data(), r->information, sizeof(r->information));
    res.error = r->error;
    res.inliers = r->inliers;
    res.valid_live_points = r->valid_live_points;
    res.valid_model_points = r->valid_model_points;
    int dofs = 0;
    if (degeneracy_rel > 0.0f) {
        const T::ObservableMotion om =
            T::keepObservableMotion(prev_pose, res.pose.data(), res.information.data(), degeneracy_rel);
        res.pose = om.pose;
        dofs = om.degenerate_dofs;
    }
    const T::TrackQuality q = T::classifyTracking(res, prev_pose, pol);
    out->quality = q == T::TrackQuality::Good ? AZU_TRACK_GOOD
                                              : (q == T::TrackQuality::Poor ? AZU_TRACK_POOR : AZU_TRACK_FAILED);
    out->degenerate_dofs = dofs;
    out->integrate = q == T::TrackQuality::Good ? 1 : 0;
    std::memcpy(out->pose, q == T::TrackQuality::Failed ? prev_pose : res.pose.data(), sizeof(out->pose));
    return AZU_OK;
}

int azu_track_weakest_ratio(const float information[36], float* out) {
    if (!information || !out) return fail(AZU_ERR_ARG, "null argument");
    *out = kfusion::tracking::weakestDirectionRatio(information);
    return AZU_OK;
}

int azu_tracker_pose_device(azu_tracker* t, const float** d_pose) {
    if (!t || !d_pose) return fail(AZU_ERR_ARG, "null argument");
    *d_pose = static_cast<const float*>(t->state);  // IcpState::pose is its first member
    return AZU_OK;
}

int azu_tracker_result(azu_tracker* t, azu_icp_result* out, azu_stream s) {
    if (!t || !out) return fail(AZU_ERR_ARG, "null argument");
    if (!t->tracked) return fail(AZU_ERR_STATE, "no track has been enqueued");
    AZU_HIP_TRY(hipMemcpyAsync(t->host_state, t->state, icpStateBytes(), hipMemcpyDeviceToHost, (hipStream_t)s),
                "state read");
    AZU_HIP_TRY(hipStreamSynchronize((hipStream_t)s), "state read sync");
    int counters[5];
    icpStateRead(t->host_state, out->pose, out->information, &out->error, &out->final_step, &out->inliers,
                 &out->converged, counters, &out->iterations);
    out->valid_live_points = counters[0];
    out->valid_model_points = counters[1];
    out->projected_points = counters[2];
    out->dist_filtered = counters[3];
    out->angle_filtered = counters[4];
    bool fin = true;
    for (float x : out->pose) fin = fin && std::isfinite(x);
    // ICPTracker::track's success rule (ICPTracker.cpp:62-64)
    out->tracking_ok = (fin && out->inliers > 100 && (out->converged || out->final_step <= 1e-3f)) ? 1 : 0;
    return AZU_OK;
}

int azu_tracker_phase_times(azu_tracker* t, double out[5], int reset) {
    if (!t || !out) return fail(AZU_ERR_ARG, "null argument");
    AZU_HIP_TRY(hipDeviceSynchronize(), "synchronise for phase times");
    AZU_HIP_TRY(hipMemcpy(t->host_state, t->state, icpStateBytes(), hipMemcpyDeviceToHost), "state read");
    unsigned long long ticks[5];
    icpTicksRead(t->host_state, ticks);
    out[0] = ticks[0] * 1e-5;  // s_memrealtime runs at 100 MHz
    out[1] = ticks[1] * 1e-5;
    out[2] = (double)ticks[2];
    out[3] = (double)ticks[3];
    out[4] = (double)ticks[4];
    if (reset) {
        AZU_HIP_TRY(launchIcpClearTicks(t->state, nullptr), "tick clear launch");
        AZU_HIP_TRY(hipDeviceSynchronize(), "tick clear");
    }
    return AZU_OK;
}

// ---- depth conditioning ----------------------------------------------------------
int azu_cond_destroy(azu_conditioner* c) {
    if (!c) return AZU_OK;
    if (c->buf[0]) (void)hipFree(c->buf[0]);
    if (c->buf[1]) (void)hipFree(c->buf[1]);
    if (c->ema) (void)hipFree(c->ema);
    delete c;
    return AZU_OK;
}

int azu_cond_create(int W, int H, azu_conditioner** out) {
    if (!out) return fail(AZU_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (W < 4 || H < 4 || W > 8192 || H > 8192) return fail(AZU_ERR_ARG, "depth frame size must be in [4, 8192]");
    auto* c = new azu_conditioner();
    c->W = W;
    c->H = H;
    const size_t n = (size_t)W * H;
    const bool ok = hipMalloc(&c->buf[0], n * 2) == hipSuccess && hipMalloc(&c->buf[1], n * 2) == hipSuccess &&
                    hipMalloc(&c->ema, n * 4) == hipSuccess;
    if (!ok) {
        azu_cond_destroy(c);
        return fail(AZU_ERR_ALLOC, "conditioner allocation failed");
    }
    if (launchCondZero(c->ema, n, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        azu_cond_destroy(c);
        return fail(AZU_ERR_HIP, "conditioner initialisation failed");
    }
    *out = c;
    return AZU_OK;
}

int azu_cond_reset_ema(azu_conditioner* c, azu_stream s) {
    if (!c) return fail(AZU_ERR_ARG, "null conditioner");
    AZU_HIP_TRY(launchCondZero(c->ema, (size_t)c->W * c->H, (hipStream_t)s), "EMA reset launch");
    return AZU_OK;
}

int azu_cond_set_minimal_depth(azu_conditioner* c, int minimal) {
    if (!c) return fail(AZU_ERR_ARG, "null conditioner");
    c->minimal = minimal != 0;
    return AZU_OK;
}

int azu_cond_run_stages(azu_conditioner* c, const uint16_t* d_in, const float* d_luma, uint16_t* d_out, float min_m,
                        float max_m, int stages, azu_stream s) {
    if (!c) return fail(AZU_ERR_ARG, "null conditioner");
    if (!d_in || !d_out) return fail(AZU_ERR_ARG, "null depth buffer");
    if (stages & ~AZU_COND_ALL) return fail(AZU_ERR_ARG, "unknown conditioning stage bit");
    if (std::isnan(min_m) || std::isnan(max_m)) return fail(AZU_ERR_ARG, "depth band must not be NaN");
    const hipStream_t st = (hipStream_t)s;
    const size_t n = (size_t)c->W * c->H;
    const int order[4] = {AZU_COND_MEDIAN, AZU_COND_HOLES, AZU_COND_GUIDED, AZU_COND_EMA};
    const char* names[4] = {"cond_median", "cond_holes", "cond_guided", "cond_ema"};
    int last = -1;
    for (int k = 0; k < 4; ++k)
        if (stages & order[k]) last = k;
    if (last < 0) {
        if (d_out != d_in) AZU_HIP_TRY(hipMemcpyAsync(d_out, d_in, n * 2, hipMemcpyDeviceToDevice, st), "depth copy");
        return AZU_OK;
    }
    // Every stage reads whole windows of its source, so no stage writes the buffer it
    // reads: an aliased caller frame is first copied into the handle's own buffers.
    const uint16_t* src = d_in;
    if (d_in == d_out) {
        AZU_HIP_TRY(hipMemcpyAsync(c->buf[1], d_in, n * 2, hipMemcpyDeviceToDevice, st), "depth copy");
        src = c->buf[1];
    }
    for (int k = 0; k <= last; ++k) {
        if (!(stages & order[k])) continue;
        uint16_t* dst = k == last ? d_out : (src == c->buf[0] ? c->buf[1] : c->buf[0]);
        ProfScope ps(names[k], st);
        AZU_HIP_TRY(launchCondStage(order[k], src, d_luma, dst, c->ema, c->W, c->H, min_m, max_m, st),
                    "conditioning launch");
        src = dst;
    }
    return AZU_OK;
}

int azu_cond_process(azu_conditioner* c, const uint16_t* d_in, const float* d_luma, uint16_t* d_out, float min_m,
                     float max_m, azu_stream s) {
    if (!c) return fail(AZU_ERR_ARG, "null conditioner");
    return azu_cond_run_stages(c, d_in, d_luma, d_out, min_m, max_m, c->minimal ? AZU_COND_MEDIAN : AZU_COND_ALL,
                               s);
}

int azu_cond_read_ema(azu_conditioner* c, float* dst, azu_stream s) {
    if (!c || !dst) return fail(AZU_ERR_ARG, "null argument");
    AZU_HIP_TRY(hipMemcpyAsync(dst, c->ema, (size_t)c->W * c->H * 4, hipMemcpyDefault, (hipStream_t)s),
                "EMA state copy");
    return AZU_OK;
}

int azu_cond_write_ema(azu_conditioner* c, const float* src, azu_stream s) {
    if (!c || !src) return fail(AZU_ERR_ARG, "null argument");
    AZU_HIP_TRY(hipMemcpyAsync(c->ema, src, (size_t)c->W * c->H * 4, hipMemcpyDefault, (hipStream_t)s),
                "EMA state copy");
    return AZU_OK;
}

// ---- profiling -------------------------------------------------------------------
int azu_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
    return AZU_OK;
}
int azu_profile_reset(void) {
    profCollect();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.clear();
    return AZU_OK;
}
int azu_profile_read(char* names, int n_max, long long* launches, double* total_ms, int* n) {
    profCollect();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    int i = 0;
    for (const auto& kv : g_prof) {
        if (i >= n_max) break;
        if (names) {
            std::memset(names + 48 * i, 0, 48);
            std::strncpy(names + 48 * i, kv.first.c_str(), 47);
        }
        if (launches) launches[i] = kv.second.launches;
        if (total_ms) total_ms[i] = kv.second.total_ms;
        ++i;
    }
    if (n) *n = i;
    return AZU_OK;
}

}  // extern "C"
