// FP64 CD-RMP2 energy that allocates nothing (mshqc_hip_rmp2_energy_noalloc).
//
// It runs the kernels of mshqc_hip_rmp2_energy's FP64 form -- the pair-block
// kernel over the pairs i <= j and 64 x 64 tiles A <= B, then the ordered
// single-workgroup sum -- with its scratch in a fixed device array instead of
// the context's workspace:
//   [tile list | pair list | pair weights | one energy partial per (pair, tile)]
// The lists are generated on the device, so no host buffer is copied either.
// The array's device address is looked up once per GPU when a context is
// created on it (mshqc_hip_ctx_create calls noalloc_init).
#include <algorithm>
#include <mutex>

#include "mp2_internal.h"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr int TILE = 64;

// One copy per GPU, held until the process ends.  Host code never uses this
// name as an address: it uses the pointer noalloc_init() looked up.
__device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_lock;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Tile list (A, B), A <= B, row-major over A (the order run_pairs builds),
// and pair list (i, j), i <= j, row-major over i, with weight 1 on the
// diagonal and 2 off it.  One thread per list entry; the decoding walks the
// rows, which are at most o (or ceil(v / 64)) long.
__global__ void noalloc_lists_kernel(int o, int ntA, int2* __restrict__ tiles, int2* __restrict__ pairs,
                                     double* __restrict__ w) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ntiles = (long)ntA * (ntA + 1) / 2;
    const long npairs = (long)o * (o + 1) / 2;
    if (k < ntiles) {
        int A = 0;
        long rem = k;
        while (rem >= ntA - A) {
            rem -= ntA - A;
            ++A;
        }
        tiles[k] = make_int2(A, A + (int)rem);
    }
    if (k < npairs) {
        int i = 0;
        long rem = k;
        while (rem >= o - i) {
            rem -= o - i;
            ++i;
        }
        pairs[k] = make_int2(i, i + (int)rem);
        w[k] = rem == 0 ? 1.0 : 2.0;
    }
}

}  // namespace

// Looks up this GPU's copy of `scratch` (the calling thread has `device`
// current).  Called by mshqc_hip_ctx_create; later calls for the same device
// return at once.
int noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return MSHQC_HIP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_lock);
    if (g_scratch[device]) return MSHQC_HIP_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        hipGetLastError();
        return MSHQC_HIP_ERR_DEVICE;
    }
    g_scratch[device] = p;
    return MSHQC_HIP_OK;
}

extern "C" int mshqc_hip_rmp2_energy_noalloc(mshqc_hip_ctx* ctx, const double* B, int Q, int o, int v,
                                             const double* eps_o, const double* eps_v, double* E) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, B && eps_o && eps_v && E, "rmp2_energy_noalloc: null pointer");
    if (Q <= 0 || o <= 0 || v <= 0)
        return set_error(ctx, MSHQC_HIP_ERR_ARG,
                         "rmp2_energy_noalloc: dimensions must be positive (Q=%d o=%d v=%d)", Q, o, v);
    if ((long)o * v > (1L << 30) || (long)Q * o * v > (1L << 40))
        return set_error(ctx, MSHQC_HIP_ERR_ARG, "rmp2_energy_noalloc: dimensions too large");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_lock);
        if (ctx->device < NOALLOC_MAX_DEVICES) base = g_scratch[ctx->device];
    }
    if (!base)
        return set_error(ctx, MSHQC_HIP_ERR_DEVICE, "rmp2_energy_noalloc: no scratch address for device %d",
                         ctx->device);
    const int ntA = (v + TILE - 1) / TILE;
    const long ntiles = (long)ntA * (ntA + 1) / 2;
    const long npairs = (long)o * (o + 1) / 2;
    const long nwork = npairs * ntiles;
    const size_t o_pairs = align256(sizeof(int2) * ntiles);
    const size_t o_w = o_pairs + align256(sizeof(int2) * npairs);
    const size_t o_part = o_w + align256(sizeof(double) * npairs);
    const size_t need = o_part + sizeof(double) * (size_t)nwork;
    if (need > NOALLOC_SCRATCH_BYTES)
        return set_error(ctx, MSHQC_HIP_ERR_ARG,
                         "rmp2_energy_noalloc: o=%d v=%d needs %.1f MiB of scratch, more than its fixed %.0f MiB "
                         "(use mshqc_hip_rmp2_energy)",
                         o, v, need / kMiB, NOALLOC_SCRATCH_BYTES / kMiB);
    char* s = static_cast<char*>(base);
    int2* tiles = reinterpret_cast<int2*>(s);
    int2* pairs = reinterpret_cast<int2*>(s + o_pairs);
    double* w = reinterpret_cast<double*>(s + o_w);
    double* part = reinterpret_cast<double*>(s + o_part);
    const long nl = std::max(ntiles, npairs);
    noalloc_lists_kernel<<<(unsigned)((nl + 255) / 256), 256, 0, ctx->stream>>>(o, ntA, tiles, pairs, w);
    LAUNCH_CHECK(ctx);
    PairArgs a{};
    a.Bl = a.Br = B;
    a.ol = a.orr = o;
    a.vl = a.vr = v;
    a.Q = Q;
    a.eol = a.eor = eps_o;
    a.evl = a.evr = eps_v;
    a.pairs = pairs;
    a.pw = w;
    a.npairs = (int)npairs;
    a.tiles = tiles;
    a.ntiles = (int)ntiles;
    a.nwork = nwork;
    a.e_part = part;
    if ((rc = launch_pair_f64(ctx, MODE_RMP2, a))) return rc;
    // every partial is written by its work item, and E is overwritten (not
    // accumulated), so no part of the scratch needs clearing between calls
    return reduce_sum(ctx, part, nwork, E, false);
}
