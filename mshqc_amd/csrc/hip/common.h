// Shared host/device definitions for the MI355X (gfx950) Cholesky-vector library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../../include/mshqc_hip.h"

// ---------------------------------------------------------------- context

// Default cap, in doubles, of the per-block amplitude workspaces (8 GiB).
constexpr size_t ZBLOCK_DEFAULT = size_t(1) << 30;

struct mshqc_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;     // stream all work is queued on
    hipStream_t own_stream = nullptr; // created by ctx_create, destroyed with the ctx
    std::string last_error;
    // Grow-only device workspace, one slab per purpose.
    struct Slab {
        void* ptr = nullptr;
        size_t bytes = 0;
    };
    Slab slabs[16];
    // Event pair for timing the last compute call (milliseconds).
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    // Zero-block sizing knob for the gradient workspace (elements of Z).
    size_t zblock_elems = ZBLOCK_DEFAULT;
    // Per-kernel-class timing.
    bool prof = false;
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double prof_ms[4] = {0, 0, 0, 0};
    long prof_n[4] = {0, 0, 0, 0};
};

enum ProfKind { PROF_PAIR = 0, PROF_GEMM = 1, PROF_J = 2 };

// Brackets one launch with events when profiling is on (no-op otherwise).
struct ProfScope {
    mshqc_hip_ctx* ctx;
    int idx = -1;
    ProfScope(mshqc_hip_ctx* c, int kind);
    ~ProfScope();
};

enum WsSlot {
    WS_PART = 0,   // per-workgroup partial sums
    WS_PART2 = 1,  // per-workgroup vector partials
    WS_Z = 2,      // pair-block amplitude workspace
    WS_MISC = 3,   // pair/tile lists, scalars
    WS_SLICE = 4,  // split-BF16 operand slices
    WS_SLICE2 = 5,
    WS_TMP = 6,    // GEMM temporaries (J/K, transforms), energy slots
    WS_TMP2 = 7,
    WS_ZS = 8,     // split-BF16 slices of the amplitude block
    WS_ZTS = 9,    // ... and of its transpose
    WS_OMP2_V = 10,  // OMP2 pair-integral block (ia|kb)
    WS_OMP2_T = 11,  // OMP2 amplitude block
    WS_MP3_A = 12,   // RMP3 operands and intermediates
    WS_MP3_B = 13,
    WS_MP3_C = 14,
    WS_MP3_D = 15,
};

// Sets ctx->last_error and returns the status code.
int set_error(mshqc_hip_ctx* ctx, int code, const char* fmt, ...);

// Returns a device pointer of at least `bytes` for `slot` (grow-only).
void* workspace(mshqc_hip_ctx* ctx, int slot, size_t bytes);

#define HIP_TRY(ctx, expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error((ctx), MSHQC_HIP_ERR_DEVICE, "%s: %s (%s:%d)", #expr,       \
                             hipGetErrorString(e_), __FILE__, __LINE__);                \
    } while (0)

#define CHECK_ARG(ctx, cond, msg)                                                       \
    do {                                                                                \
        if (!(cond)) return set_error((ctx), MSHQC_HIP_ERR_ARG, "%s", msg);              \
    } while (0)

// Launch-and-check helper: checks the launch error right after queueing.
#define LAUNCH_CHECK(ctx)                                                               \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return set_error((ctx), MSHQC_HIP_ERR_DEVICE, "kernel launch: %s (%s:%d)",   \
                             hipGetErrorString(e_), __FILE__, __LINE__);                \
    } while (0)

// Enter the ctx's device; every entry point calls this first.
int ctx_enter(mshqc_hip_ctx* ctx);
// Looks up, once per GPU, the device address of the fixed scratch array of
// mshqc_hip_rmp2_energy_noalloc (noalloc.hip); `device` must be current.
int noalloc_init(int device);
// Synchronise the ctx stream and turn asynchronous faults into a status.
int ctx_finish(mshqc_hip_ctx* ctx);

// ------------------------------------------------------------- GEMM engine

// C_b[m][n] (+)= alpha * sum_k A_b[m][k] X_b[k][n] over a batch of b.
//   A_b[m][k] = transA ? A[b*sA + k*lda + m] : A[b*sA + m*lda + k]
//   X_b[k][n] = transX ? X[b*sX + n*ldx + k] : X[b*sX + k*ldx + n]
// reduce_batch = 0: one output per b (C + b*sC).
// reduce_batch = 1: sum over b into one output; the batch is split over
//   `splits` workgroup layers whose partials land in `part`
//   ([splits][M][N]) and are then summed in split order (deterministic).
// inner > 1 (reduce_batch only): batch entry b = bo * inner + bi starts at
//   A + bo*sA + bi*sA2 and X + bo*sX + bi*sX2, so one contraction can be cut
//   into chunks along a second stride.
struct GemmF64 {
    const double* A = nullptr;
    long lda = 0, sA = 0;
    bool transA = false;
    const double* X = nullptr;
    long ldx = 0, sX = 0;
    bool transX = false;
    double* C = nullptr;
    long ldc = 0, sC = 0;
    int M = 0, N = 0, K = 0, batch = 1;
    double alpha = 1.0;
    bool accumulate = false;  // C += ... instead of C = ...
    bool reduce_batch = false;
    int splits = 1;
    int inner = 1;
    long sA2 = 0, sX2 = 0;
    // M == N and C symmetric: only the 64 x 64 tiles on and below the
    // diagonal are formed (mirror_upper then fills the rest)
    bool lower_only = false;
    // each batch entry's C stored transposed: element (m, n) at
    // C + b*sC + n*ldc + m (ldc >= M); plain batches only
    bool trans_c = false;
    int tblock = 0;  // set by gemm_f64: rows per transposed block (the entry's M)
};
int gemm_f64(mshqc_hip_ctx* ctx, const GemmF64& g);

// J = sum_Q tr(L^Q P_tot) L^Q and K[s] = sum_Q L^Q P_s L^Q queued on the ctx
// stream (the work of mshqc_hip_jk without its entry checks and timing).
int jk_general(mshqc_hip_ctx* ctx, const double* L, int Q, int n, const double* P, int nspin, int screen,
               double* J, double* K);

// Square n x n matrix (row stride ld): the elements of every tile
// (64 * tpb rows and columns per block) whose column block exceeds its row
// block are overwritten by their transposes, so a lower-block-triangle
// product becomes the whole symmetric matrix.
int mirror_upper(mshqc_hip_ctx* ctx, double* M, long ld, long n, long tpb);

// Deterministic single-workgroup sum of x[0..n) into *out (device), using a
// fixed per-thread stride and a fixed tree; Neumaier compensation per thread.
int reduce_sum(mshqc_hip_ctx* ctx, const double* x, long n, double* out, bool accumulate);

// x[0..count) = 0 with a fill kernel on the ctx stream (no memset node, so
// it stays correct inside captured graphs).
int fill_zero(mshqc_hip_ctx* ctx, double* x, long count);

// Number of bytes in a mebibyte, for messages.
static constexpr double kMiB = 1024.0 * 1024.0;
