// gfx950 kernels of the parallel-in-time LTI scan.
//
// The scan is the reference's Blelloch up/down-sweep over the affine step maps
// (Sources/GPUFixedStepIterative.cpp:236-398) after the LTI collapse: every
// per-element matrix at level l equals P_l = M^(2^l), so a level is one
// GEMM-shaped update  X[R] += P_l X[L]  (up) or  (X[L], X[R]) <- (X[R],
// P_l X[R] + X[L])  (down) over the strided columns of the trajectory X, and
// the reference's MM kernels, left-leaf and root copies collapse into the
// power table and into the in-place pair update.
#include "kernels.h"

#include <algorithm>
#include <vector>
#include <cstring>
#include <type_traits>
#include <stdexcept>
#include <string>

namespace parode {
namespace k {

// The block a kernel body works on: (x, y) and the grid's width in x.  A
// single-system launch passes its own blockIdx and gridDim; a grouped launch
// (many systems in one grid) passes the block's place in its system's grid.
struct VBlock {
    long x, y, gx;
};

static void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " +
                                 hipGetErrorString(e));
}

static inline unsigned blocks_for(long work, int block, long cap = 1L << 20)
{
    long b = (work + block - 1) / block;
    if (b < 1) b = 1;
    return (unsigned)std::min(b, cap);
}

// ---------------------------------------------------------------- helpers
template <typename T>
__device__ __forceinline__ void d_transpose(const VBlock vb, int rows, int cols, const T* __restrict__ in, T* __restrict__ out)
{
    const long total = (long)rows * cols;
    for (long idx = (long)vb.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)vb.gx * blockDim.x) {
        const long c = idx / rows, r = idx % rows;   // out[c][r] = in[r][c]
        out[idx] = in[r * cols + c];
    }
}

template <typename T>
__global__ void k_transpose(int rows, int cols, const T* __restrict__ in, T* __restrict__ out)
{
    d_transpose<T>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, rows, cols, in, out);
}

template <typename T>
void transpose(hipStream_t s, int rows, int cols, const T* in, T* out)
{
    const long work = (long)rows * cols;
    if (work <= 0) return;
    hipLaunchKernelGGL((k_transpose<T>), dim3(blocks_for(work, 256, 4096)), dim3(256), 0, s, rows,
                       cols, in, out);
    check_launch("transpose");
}

template void transpose<float>(hipStream_t, int, int, const float*, float*);
template void transpose<double>(hipStream_t, int, int, const double*, double*);

// ------------------------------------------------------------------ GEMMs
__global__ void k_gemm_seq_f32(int r, int kk, int c, const float* __restrict__ X,
                               const float* __restrict__ Y, float* __restrict__ Z)
{
#pragma clang fp contract(off)
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)r * c) return;
    int i = (int)(idx / c), j = (int)(idx % c);
    float t = 0.0f;
    for (int q = 0; q < kk; ++q) {
        float pr = X[(long)i * kk + q] * Y[(long)q * c + j];
        t = t + pr;
    }
    Z[idx] = t;
}

void gemm_seq_f32(hipStream_t s, int r, int kk, int c, const float* X, const float* Y, float* Z)
{
    long work = (long)r * c;
    hipLaunchKernelGGL(k_gemm_seq_f32, dim3(blocks_for(work, 256)), dim3(256), 0, s, r, kk, c, X,
                       Y, Z);
    check_launch("gemm_seq_f32");
}

// 64x64 output tile, 16x16 threads with 4x4 micro-tiles, K-step 16, FP64 FMA.
__global__ __launch_bounds__(256) void k_gemm_f64(int r, int kk, int c,
                                                  const double* __restrict__ X,
                                                  const double* __restrict__ Y,
                                                  double* __restrict__ Z)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int q0 = 0; q0 < kk; q0 += BK) {
        for (int t = threadIdx.x; t < BM * BK; t += 256) {
            int i = t / BK, q = t % BK;
            As[q][i] = (i0 + i < r && q0 + q < kk) ? X[(long)(i0 + i) * kk + q0 + q] : 0.0;
        }
        for (int t = threadIdx.x; t < BK * BN; t += 256) {
            int q = t / BN, j = t % BN;
            Bs[q][j] = (q0 + q < kk && j0 + j < c) ? Y[(long)(q0 + q) * c + j0 + j] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BK; ++q) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[q][ty + 16 * u];
                b[u] = Bs[q][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
            if (i < r && j < c) Z[(long)i * c + j] = acc[u][v];
        }
}

__global__ void k_gemm_f64_mfma(int r, int kk, int c, const double* __restrict__ X,
                                const double* __restrict__ Y, double* __restrict__ Z);

// FP64 GEMM for the squarings of wide power tables: 32 x 32 tiles of Z (so
// n_a = 584 .. 1024 gives 361 .. 1024 work-groups), K split among KSPLIT
// groups of four waves, each wave one 16 x 16 quadrant over every KSPLIT-th
// k-step of a 32-deep slice staged in LDS; the groups' partial tiles meet in
// LDS after the last slice and group 0 adds them in group order.  More waves
// per SIMD than one-wave-per-tile kernels, whose dependent FP64 MFMAs leave the
// pipe idle.  Loads are bounds-checked; the next slice is in flight while the
// current one is multiplied.
typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kGemmKsT = 32, kGemmKsKB = 32;

// Leading dimensions of the FP64 GEMMs' operands (X r x kk, Y kk x c, Z r x c,
// row-major), and an optional addend: Z[i][j] = (X Y)[i][j] + add[(i - shift)
// ldz + j] for i >= shift (the ABM companion's shifted blocks).
struct GemmLd {
    int ldx, ldy, ldz;
    const double* add;
    int shift;
};
__host__ __device__ inline GemmLd gemm_ld(int kk, int c) { return GemmLd{kk, c, c, nullptr, 0}; }
__device__ __forceinline__ double gemm_out(const GemmLd& ld, int i, int j, double v)
{
    return (ld.add && i >= ld.shift) ? v + ld.add[(long)(i - ld.shift) * ld.ldz + j] : v;
}

template <int KSPLIT>
__device__ __forceinline__ void d_gemm_f64_ks(const VBlock vb, int r, int kk, int c,
                                                              const double* __restrict__ X,
                                                              const double* __restrict__ Y,
                                                              double* __restrict__ Z,
                                                              const GemmLd ld)
{
    constexpr int T = kGemmKsT, KB = kGemmKsKB, AP = KB + 1, BP = T + 1, NT = 256 * KSPLIT;
    constexpr int PA = T * KB / NT, PB = KB * T / NT;
    static_assert(PA >= 1 && PB >= 1 && 4 * KSPLIT <= KB, "tile, slice and split");
    extern __shared__ double gsm[];
    double* As = gsm;                      // [T][AP]
    double* Bs = gsm + T * AP;             // [KB][BP]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int grp = wave >> 2, wq = wave & 3;
    const int wr = wq >> 1, wc = wq & 1, l16 = lane & 15, q4 = lane >> 4;
    const int i0 = vb.y * T, j0 = vb.x * T;
    const int ar = tid / (KB / PA), ak = (tid % (KB / PA)) * PA;
    const int bk = tid / (T / PB), bj = (tid % (T / PB)) * PB;
    double ra[PA], rb[PB];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int gi = i0 + ar, gk = k0 + ak + q;
            ra[q] = (gi < r && gk < kk) ? X[(long)gi * ld.ldx + gk] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int gk2 = k0 + bk, gj = j0 + bj + q;
            rb[q] = (gk2 < kk && gj < c) ? Y[(long)gk2 * ld.ldy + gj] : 0.0;
        }
    };
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    load(0);
    for (int k0 = 0; k0 < kk; k0 += KB) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PA; ++q) As[ar * AP + ak + q] = ra[q];
#pragma unroll
        for (int q = 0; q < PB; ++q) Bs[bk * BP + bj + q] = rb[q];
        __syncthreads();
        if (k0 + KB < kk) load(k0 + KB);
#pragma unroll
        for (int k4 = 4 * grp; k4 < KB; k4 += 4 * KSPLIT) {
            const double a = As[(wr * 16 + l16) * AP + k4 + q4];
            const double b = Bs[(k4 + q4) * BP + wc * 16 + l16];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
    }
    __syncthreads();                      // the staging area holds the partial tiles now
    double* part = gsm;                   // [KSPLIT - 1][4 waves][4 values][64 lanes]
    if (grp > 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) part[(((grp - 1) * 4 + wq) * 4 + v) * 64 + lane] = acc[v];
    }
    __syncthreads();
    if (grp > 0) return;
    // D[i][j] is in lane j + 16 (i % 4), register i / 4
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double y = acc[v];
#pragma unroll
        for (int g = 1; g < KSPLIT; ++g) y += part[(((g - 1) * 4 + wq) * 4 + v) * 64 + lane];
        const int i = i0 + wr * 16 + 4 * v + q4, j = j0 + wc * 16 + l16;
        if (i < r && j < c) Z[(long)i * ld.ldz + j] = gemm_out(ld, i, j, y);
    }
}

template <int KSPLIT>
__global__ __launch_bounds__(256 * KSPLIT) void k_gemm_f64_ks(int r, int kk, int c,
                                                              const double* __restrict__ X,
                                                              const double* __restrict__ Y,
                                                              double* __restrict__ Z)
{
    d_gemm_f64_ks<KSPLIT>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, r, kk, c, X, Y, Z,
                          gemm_ld(kk, c));
}

// FP64 GEMM on 64 x 64 tiles of Z for the squarings of the power tables: four
// waves, each a 32 x 32 quadrant as 2 x 2 tiles of v_mfma_f64_16x16x4f64 (four
// independent accumulators, each operand fragment used twice), K in slices of
// 16 staged in LDS (odd pitches), double buffered: the next slice's loads are
// in flight while the current one is multiplied, one barrier per slice.  Twice
// the arithmetic per byte of operands of the 32 x 32 tiles.
constexpr int kG64 = 64, kG64KB = 16;

// 64 x 64 tiles once the padding of the edge tiles costs less than the 16 x 16
// kernel's lower rate (about half)
__host__ __device__ constexpr bool gemm_wide(int r, int kk, int c)
{
    return r >= 96 && c >= 96 && kk >= 64;
}

__device__ __forceinline__ void d_gemm_f64_t64(const VBlock vb, int r, int kk, int c,
                                               const double* __restrict__ X,
                                               const double* __restrict__ Y,
                                               double* __restrict__ Z, const GemmLd ld)
{
    __shared__ double As[2][kG64][kG64KB + 1];
    __shared__ double Bs[2][kG64KB][kG64 + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1, l16 = lane & 15, q4 = lane >> 4;
    const int i0 = (int)vb.y * kG64, j0 = (int)vb.x * kG64;
    const int ar = tid >> 2, ak = (tid & 3) * 4;     // A slice: row, first of four k
    const int bk = tid >> 4, bj = (tid & 15) * 4;    // B slice: k, first of four columns
    double ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int gi = i0 + ar, gk = k0 + ak + q;
            ra[q] = (gi < r && gk < kk) ? X[(long)gi * ld.ldx + gk] : 0.0;
            const int gk2 = k0 + bk, gj = j0 + bj + q;
            rb[q] = (gk2 < kk && gj < c) ? Y[(long)gk2 * ld.ldy + gj] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            As[buf][ar][ak + q] = ra[q];
            Bs[buf][bk][bj + q] = rb[q];
        }
    };
    doublex4 acc[2][2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = doublex4{0.0, 0.0, 0.0, 0.0};
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < kk; k0 += kG64KB) {
        const bool more = k0 + kG64KB < kk;
        if (more) load(k0 + kG64KB);
#pragma unroll
        for (int ks = 0; ks < kG64KB; ks += 4) {
            double a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = As[buf][wr * 32 + t * 16 + l16][ks + q4];
                b[t] = Bs[buf][ks + q4][wc * 32 + t * 16 + l16];
            }
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj],
                                                                      0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // D[i][j] is in lane j + 16 (i % 4), register i / 4
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wr * 32 + ti * 16 + 4 * v + q4;
                const int j = j0 + wc * 32 + tj * 16 + l16;
                if (i < r && j < c) Z[(long)i * ld.ldz + j] = gemm_out(ld, i, j, acc[ti][tj][v]);
            }
}

void gemm_f64(hipStream_t s, int r, int kk, int c, const double* X, const double* Y, double* Z,
              bool mfma)
{
    if (mfma && r >= 320 && c >= 320 && kk >= 64) {
        const dim3 grid((c + kGemmKsT - 1) / kGemmKsT, (r + kGemmKsT - 1) / kGemmKsT);
        const size_t stage = (size_t)(kGemmKsT * (kGemmKsKB + 1) + kGemmKsKB * (kGemmKsT + 1));
        // four K groups while the tiles are few, two from 512 tiles on
        if ((long)grid.x * grid.y < 512) {
            const size_t lds = std::max<size_t>(stage, 3 * 4 * 4 * 64) * sizeof(double);
            hipLaunchKernelGGL((k_gemm_f64_ks<4>), grid, dim3(1024), lds, s, r, kk, c, X, Y, Z);
        } else {
            const size_t lds = std::max<size_t>(stage, 1 * 4 * 4 * 64) * sizeof(double);
            hipLaunchKernelGGL((k_gemm_f64_ks<2>), grid, dim3(512), lds, s, r, kk, c, X, Y, Z);
        }
    } else if (mfma) {
        dim3 grid((c + 15) / 16, (r + 15) / 16);
        hipLaunchKernelGGL(k_gemm_f64_mfma, grid, dim3(256), 0, s, r, kk, c, X, Y, Z);
    } else {
        dim3 grid((c + 63) / 64, (r + 63) / 64);
        hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, s, r, kk, c, X, Y, Z);
    }
    check_launch("gemm_f64");
}

// FP64 GEMM on the matrix cores (v_mfma_f64_16x16x4_f64), Z = X Y, for the
// small products of the step map and the squarings of the power table.  A
// work-group of four waves owns one 16x16 tile of Z and the waves split the K
// range in four: each wave issues the loads of 16 k-steps before their MFMAs,
// so a 256-deep product waits on one round of loads, not one per k-slice.  The
// four partial tiles are summed through LDS.  Operands come from L2 (the
// matrices are at most a few MiB).

__device__ __forceinline__ void d_gemm_f64_mfma(const VBlock vb, int r, int kk, int c,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ Y,
                                                       double* __restrict__ Z, const GemmLd ld)
{
    __shared__ double part[3][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
    const int i0 = vb.y * 16, j0 = vb.x * 16;
    const int ksteps = (kk + 3) / 4;
    const int per = (ksteps + 3) / 4;
    const int kb = wave * per, ke = (kb + per) < ksteps ? kb + per : ksteps;
    const int ia = i0 + l16, jb = j0 + l16;
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    constexpr int U = 16;
    for (int k = kb; k < ke; k += U) {
        double a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kx = (k + u) * 4 + q4;
            const bool ok = k + u < ke && kx < kk;
            a[u] = (ok && ia < r) ? X[(long)ia * ld.ldx + kx] : 0.0;
            b[u] = (ok && jb < c) ? Y[(long)kx * ld.ldy + jb] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    if (wave > 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) part[wave - 1][v][lane] = acc[v];
    }
    __syncthreads();
    if (wave == 0) {
        // v_mfma_f64_16x16x4_f64 returns D[i][j] in lane j + 16 (i % 4), register i / 4
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double sum = acc[v] + part[0][v][lane] + part[1][v][lane] + part[2][v][lane];
            const int i = i0 + 4 * v + q4, j = j0 + l16;
            if (i < r && j < c) Z[(long)i * ld.ldz + j] = gemm_out(ld, i, j, sum);
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm_f64_mfma(int r, int kk, int c,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ Y,
                                                       double* __restrict__ Z)
{
    d_gemm_f64_mfma(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, r, kk, c, X, Y, Z,
                    gemm_ld(kk, c));
}

__device__ __forceinline__ void d_cast_f32_f64(const VBlock vb, long count, const float* __restrict__ in, double* __restrict__ out)
{
    for (long i = (long)vb.x * blockDim.x + threadIdx.x; i < count;
         i += (long)vb.gx * blockDim.x)
        out[i] = (double)in[i];
}

__global__ void k_cast_f32_f64(long count, const float* __restrict__ in, double* __restrict__ out)
{
    d_cast_f32_f64(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, count, in, out);
}

void cast_f32_to_f64(hipStream_t s, long count, const float* in, double* out)
{
    hipLaunchKernelGGL(k_cast_f32_f64, dim3(blocks_for(count, 256, 4096)), dim3(256), 0, s, count,
                       in, out);
    check_launch("cast_f32_to_f64");
}

// Zero/constant fill as a kernel (graph-capturable; no memset nodes).
template <typename T>
__global__ void k_fill(long count, T* __restrict__ out)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long)gridDim.x * blockDim.x)
        out[i] = (T)0;
}

__device__ __forceinline__ void d_fill_f64(const VBlock vb, long count, double v, double* __restrict__ out)
{
    for (long i = (long)vb.x * blockDim.x + threadIdx.x; i < count;
         i += (long)vb.gx * blockDim.x)
        out[i] = v;
}

__global__ void k_fill_f64(long count, double v, double* __restrict__ out)
{
    d_fill_f64(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, count, v, out);
}

void fill_zero_f32(hipStream_t s, long count, float* out)
{
    if (count <= 0) return;
    hipLaunchKernelGGL((k_fill<float>), dim3(blocks_for(count, 256, 8192)), dim3(256), 0, s, count,
                       out);
    check_launch("fill_zero_f32");
}

void fill_f64(hipStream_t s, long count, double v, double* out)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(count, 256, 4096)), dim3(256), 0, s, count, v,
                       out);
    check_launch("fill_f64");
}

// ------------------------------------------------------------- step maps
// Scalar factors evaluated left to right as in RungeKutta4thOrder.cpp:48-57.
static Rk4Coef rk4_coef(double h)
{
#pragma clang fp contract(off)
    Rk4Coef c;
    c.h = h;
    c.c2 = 0.5 * h * h;
    c.c3 = (1.0 / 6.0) * h * h * h;
    c.c4 = (1.0 / 24.0) * h * h * h * h;
    c.s1 = h * h;
    c.s2 = 0.5 * h * h * h;
    c.s3 = 0.25 * h * h * h * h;
    c.t0 = 4.0 * h;
    c.t1 = 2.0 * h * h;
    c.t2 = 0.5 * h * h * h;
    c.u0 = (1.0 / 6.0) * h;
    return c;
}

Rk4Coef rk4_coefficients(double h) { return rk4_coef(h); }

template <typename TI>
__device__ __forceinline__ void d_rk4_combine(const VBlock vb, bool mirror, int n, int m, Rk4Coef k, const TI* __restrict__ A,
                              const TI* __restrict__ A2, const TI* __restrict__ A3,
                              const TI* __restrict__ A4, const TI* __restrict__ B,
                              const TI* __restrict__ AB, const TI* __restrict__ A2B,
                              const TI* __restrict__ A3B, double* __restrict__ MT64,
                              double* __restrict__ Bbar64, float* __restrict__ Bbar32)
{
#pragma clang fp contract(off)
    long idx = (long)vb.x * blockDim.x + threadIdx.x;
    long nn = (long)n * n;
    if (idx < nn) {
        int i = (int)(idx / n), j = (int)(idx % n);
        double v = ((i == j) ? 1.0 : 0.0) + k.h * (double)A[idx];
        v = v + k.c2 * (double)A2[idx];
        v = v + k.c3 * (double)A3[idx];
        v = v + k.c4 * (double)A4[idx];
        if (mirror) v = (double)(float)v;
        MT64[(long)j * n + i] = v;                 // the power table is kept transposed
        return;
    }
    idx -= nn;
    if (idx >= (long)n * m) return;
    int i = (int)(idx / m), j = (int)(idx % m);
    double b = B[idx], ab = AB[idx], a2b = A2B[idx], a3b = A3B[idx];
    double b1 = (1.0 / 6.0) * (((k.h * b + k.s1 * ab) + k.s2 * a2b) + k.s3 * a3b);
    double b2 = (1.0 / 6.0) * ((k.t0 * b + k.t1 * ab) + k.t2 * a2b);
    double b3 = k.u0 * b;
    double vals[3] = {b1, b2, b3};
    for (int q = 0; q < 3; ++q) {
        long o = (long)i * 3 * m + q * m + j;
        float f = (float)vals[q];
        if (Bbar32) Bbar32[o] = f;
        Bbar64[o] = mirror ? (double)f : vals[q];
    }
}

template <typename TI>
__global__ void k_rk4_combine(bool mirror, int n, int m, Rk4Coef k, const TI* __restrict__ A,
                              const TI* __restrict__ A2, const TI* __restrict__ A3,
                              const TI* __restrict__ A4, const TI* __restrict__ B,
                              const TI* __restrict__ AB, const TI* __restrict__ A2B,
                              const TI* __restrict__ A3B, double* __restrict__ MT64,
                              double* __restrict__ Bbar64, float* __restrict__ Bbar32)
{
    d_rk4_combine<TI>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, mirror, n, m, k, A, A2, A3, A4, B, AB, A2B, A3B, MT64, Bbar64, Bbar32);
}

void rk4_combine(hipStream_t s, bool mirror, int n, int m, double h, const void* A, const void* A2,
                 const void* A3, const void* A4, const void* B, const void* AB, const void* A2B,
                 const void* A3B, double* M64, double* Bbar64, float* Bbar32)
{
    Rk4Coef k = rk4_coef(h);
    long work = (long)n * n + (long)n * m;
    if (mirror)
        hipLaunchKernelGGL(k_rk4_combine<float>, dim3(blocks_for(work, 256)), dim3(256), 0, s, true,
                           n, m, k, (const float*)A, (const float*)A2, (const float*)A3,
                           (const float*)A4, (const float*)B, (const float*)AB, (const float*)A2B,
                           (const float*)A3B, M64, Bbar64, Bbar32);
    else
        hipLaunchKernelGGL(k_rk4_combine<double>, dim3(blocks_for(work, 256)), dim3(256), 0, s,
                           false, n, m, k, (const double*)A, (const double*)A2, (const double*)A3,
                           (const double*)A4, (const double*)B, (const double*)AB,
                           (const double*)A2B, (const double*)A3B, M64, Bbar64, Bbar32);
    check_launch("rk4_combine");
}

// AdamsMoulton.cpp:258-269, each factor evaluated left to right.
static AbmCoef abm_coef(double h)
{
#pragma clang fp contract(off)
    AbmCoef c;
    c.a0A = 28.0 * h / 24.0;
    c.a0A2 = 495.0 * h * h / 576.0;
    c.a1A = -5.0 * h / 24.0;
    c.a1A2 = -531.0 * h * h / 576.0;
    c.a2A = 1.0 * h / 24.0;
    c.a2A2 = 333.0 * h * h / 576.0;
    c.a3A2 = -81.0 * h * h / 576.0;
    c.bm1 = 9.0 * h / 24.0;
    c.b0B = 19.0 * h / 24.0;
    c.b0AB = 495.0 * h * h / 576.0;
    c.b1B = -5.0 * h / 24.0;
    c.b1AB = -531.0 * h * h / 576.0;
    c.b2B = h / 24.0;
    c.b2AB = 333.0 * h * h / 576.0;
    c.b3AB = -81.0 * h * h / 576.0;
    return c;
}

AbmCoef abm_coefficients(double h) { return abm_coef(h); }

template <typename TI>
__device__ __forceinline__ void d_abm_combine(const VBlock vb, bool mirror, int n, int m, AbmCoef c, const TI* __restrict__ A,
                              const TI* __restrict__ A2, const TI* __restrict__ B,
                              const TI* __restrict__ AB, double* __restrict__ MT64,
                              double* __restrict__ Bbar64, float* __restrict__ Bbar32)
{
#pragma clang fp contract(off)
    long idx = (long)vb.x * blockDim.x + threadIdx.x;
    long nn = (long)n * n;
    const long na = 4L * n;
    auto rnd = [mirror](double v) { return mirror ? (double)(float)v : v; };
    if (idx < nn) {
        int r = (int)(idx / n), q = (int)(idx % n);
        double id = (r == q) ? 1.0 : 0.0, a = A[idx], a2 = A2[idx];
        double blk[4] = {rnd(c.a3A2 * a2), rnd(c.a2A * a + c.a2A2 * a2),
                         rnd(c.a1A * a + c.a1A2 * a2), rnd((id + c.a0A * a) + c.a0A2 * a2)};
        const long row = 3L * n + r;
        for (int b = 0; b < 4; ++b) MT64[(long)(b * n + q) * na + row] = blk[b];   // transposed
        if (r == q) {
            MT64[(long)(n + r) * na + r] = 1.0;              // M[r][n+r]
            MT64[(long)(2 * n + r) * na + n + r] = 1.0;      // M[n+r][2n+r]
            MT64[(long)(3 * n + r) * na + 2 * n + r] = 1.0;  // M[2n+r][3n+r]
        }
        return;
    }
    idx -= nn;
    if (idx >= (long)n * m) return;
    int r = (int)(idx / m), q = (int)(idx % m);
    double b = B[idx], ab = AB[idx];
    double vals[5] = {c.b3AB * ab, c.b2B * b + c.b2AB * ab, c.b1B * b + c.b1AB * ab,
                      c.b0B * b + c.b0AB * ab, c.bm1 * b};
    for (int k = 0; k < 5; ++k) {
        long o = (long)r * 5 * m + (long)k * m + q;
        float f = (float)vals[k];
        if (Bbar32) Bbar32[o] = f;
        Bbar64[o] = mirror ? (double)f : vals[k];
    }
}

template <typename TI>
__global__ void k_abm_combine(bool mirror, int n, int m, AbmCoef c, const TI* __restrict__ A,
                              const TI* __restrict__ A2, const TI* __restrict__ B,
                              const TI* __restrict__ AB, double* __restrict__ MT64,
                              double* __restrict__ Bbar64, float* __restrict__ Bbar32)
{
    d_abm_combine<TI>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, mirror, n, m, c, A, A2, B, AB, MT64, Bbar64, Bbar32);
}

void abm_combine(hipStream_t s, bool mirror, int n, int m, double h, const void* A, const void* A2,
                 const void* B, const void* AB, double* M64, double* Bbar64, float* Bbar32)
{
    AbmCoef c = abm_coef(h);
    long na = 4L * n;
    fill_f64(s, na * na, 0.0, M64);
    long work = (long)n * n + (long)n * m;
    if (mirror)
        hipLaunchKernelGGL(k_abm_combine<float>, dim3(blocks_for(work, 256)), dim3(256), 0, s, true,
                           n, m, c, (const float*)A, (const float*)A2, (const float*)B,
                           (const float*)AB, M64, Bbar64, Bbar32);
    else
        hipLaunchKernelGGL(k_abm_combine<double>, dim3(blocks_for(work, 256)), dim3(256), 0, s,
                           false, n, m, c, (const double*)A, (const double*)A2, (const double*)B,
                           (const double*)AB, M64, Bbar64, Bbar32);
    check_launch("abm_combine");
}

// -------------------------------------------------------------- level tables
template <typename T>
__global__ void k_level_table(long nn, int na, const double* __restrict__ PT64, bool delta,
                              T* __restrict__ PT)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
         i += (long)gridDim.x * blockDim.x) {
        double v = PT64[i];
        if (delta && (i / na) == (i % na)) v -= 1.0;
        PT[i] = (T)v;
    }
}

// Zero-padded FP32 table for the MFMA levels: PTp[j][i] for j < nrp, i < nap.
__global__ void k_level_table_padded(int na, int nap, int nrp, const double* __restrict__ PT64,
                                     bool delta, float* __restrict__ PTp)
{
    const long total = (long)nrp * nap;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const int j = (int)(idx / nap), i = (int)(idx % nap);
        float v = 0.0f;
        if (j < na && i < na) {
            double x = PT64[(long)j * na + i];
            if (delta && i == j) x -= 1.0;
            v = (float)x;
        }
        PTp[idx] = v;
    }
}

void make_level_table_padded(hipStream_t s, int na, int nap, int nrp, const double* P, bool delta,
                             float* PTp)
{
    const long work = (long)nrp * nap;
    hipLaunchKernelGGL(k_level_table_padded, dim3(blocks_for(work, 256, 8192)), dim3(256), 0, s, na,
                       nap, nrp, P, delta, PTp);
    check_launch("level_table_padded");
}

void make_level_table_f32(hipStream_t s, int na, const double* P, bool delta, float* PT)
{
    long nn = (long)na * na;
    hipLaunchKernelGGL(k_level_table<float>, dim3(blocks_for(nn, 256, 8192)), dim3(256), 0, s, nn,
                       na, P, delta, PT);
    check_launch("level_table_f32");
}

void make_level_table_f64(hipStream_t s, int na, const double* P, bool delta, double* PT)
{
    long nn = (long)na * na;
    hipLaunchKernelGGL(k_level_table<double>, dim3(blocks_for(nn, 256, 8192)), dim3(256), 0, s, nn,
                       na, P, delta, PT);
    check_launch("level_table_f64");
}

// All FP32 level tables of a solve in one launch: levels [0, nlev) of the
// transposed FP64 powers (delta form below deltaBelow) as the unpadded FP32
// table PT32 [na][na] (if not null) and the zero-padded MFMA table PTm
// [nrp][nap] (if not null; else nap = nrp = na).
__device__ __forceinline__ void d_level_tables(const VBlock vb, int na, int nap, int nrp, int nlev, int deltaBelow,
                               const double* __restrict__ PT64, float* __restrict__ PT32,
                               float* __restrict__ PTm)
{
    const long per = (long)nrp * nap, nn = (long)na * na;
    const long total = per * nlev;
    for (long idx = (long)vb.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)vb.gx * blockDim.x) {
        const int l = (int)(idx / per);
        const long r = idx - l * per;
        const int j = (int)(r / nap), i = (int)(r - (long)j * nap);
        float v = 0.0f;
        if (j < na && i < na) {
            double x = PT64[l * nn + (long)j * na + i];
            if (l < deltaBelow && i == j) x -= 1.0;
            v = (float)x;
            if (PT32) PT32[l * nn + (long)j * na + i] = v;
        }
        if (PTm) PTm[idx] = v;
    }
}

__global__ void k_level_tables(int na, int nap, int nrp, int nlev, int deltaBelow,
                               const double* __restrict__ PT64, float* __restrict__ PT32,
                               float* __restrict__ PTm)
{
    d_level_tables(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, nap, nrp, nlev, deltaBelow, PT64, PT32, PTm);
}

void make_level_tables(hipStream_t s, int na, int nap, int nrp, int nlev, int deltaBelow,
                       const double* PT64, float* PT32, float* PTm)
{
    if (nlev <= 0) return;
    if (!PTm) nap = nrp = na;
    const long total = (long)nrp * nap * nlev;
    hipLaunchKernelGGL(k_level_tables, dim3(blocks_for(total, 256, 16384)), dim3(256), 0, s, na,
                       nap, nrp, nlev, deltaBelow, PT64, PT32, PTm);
    check_launch("level_tables");
}

// All squarings of a small power table in one work-group: the transposed
// FP64 table P_l (n_a <= 80) sits in LDS (pitch odd in doubles); each of the
// 16 waves forms 16 x 16 tiles of P_l^2 on the FP64 matrix cores, the tiles go
// to LDS and to PT[l + 1] after a barrier.  Replaces L launches of the GEMM.
constexpr int kPowersSmallMax = 80;

__device__ __forceinline__ void d_powers_small(const VBlock vb, int na, int L, double* __restrict__ PT)
{
    extern __shared__ double Ps[];
    const int n16 = (na + 15) / 16 * 16, pitch = n16 + 1;
    const int nt = n16 / 16, ntiles = nt * nt;
    const long nn = (long)na * na;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
    for (int t = threadIdx.x; t < n16 * n16; t += 1024) {
        const int r = t / n16, c = t - r * n16;
        Ps[r * pitch + c] = (r < na && c < na) ? PT[(long)r * na + c] : 0.0;
    }
    __syncthreads();
    for (int l = 0; l < L; ++l) {
        doublex4 acc[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            acc[q] = {0.0, 0.0, 0.0, 0.0};
            const int tile = wave + 16 * q;
            if (tile >= ntiles) continue;
            const int ti = tile / nt, tj = tile - ti * nt;
            for (int k = 0; k < n16; k += 4) {
                const double a = Ps[(ti * 16 + l16) * pitch + k + q4];
                const double b = Ps[(k + q4) * pitch + tj * 16 + l16];
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
        }
        __syncthreads();
        double* out = PT + (long)(l + 1) * nn;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int tile = wave + 16 * q;
            if (tile >= ntiles) continue;
            const int ti = tile / nt, tj = tile - ti * nt;
            // D[i][j] is in lane j + 16 (i % 4), register i / 4
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ti * 16 + 4 * v + q4, c = tj * 16 + l16;
                Ps[r * pitch + c] = acc[q][v];
                if (r < na && c < na) out[(long)r * na + c] = acc[q][v];
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_powers_small(int na, int L, double* __restrict__ PT)
{
    d_powers_small(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, L, PT);
}

bool powers_small_supported(int na) { return na <= kPowersSmallMax; }

void powers_small(hipStream_t s, int na, int L, double* PT)
{
    if (L <= 0) return;
    if (na > kPowersSmallMax) throw std::invalid_argument("powers_small: n_a too large");
    const int n16 = (na + 15) / 16 * 16;
    const size_t lds = (size_t)n16 * (n16 + 1) * sizeof(double);
    hipLaunchKernelGGL(k_powers_small, dim3(1), dim3(1024), lds, s, na, L, PT);
    check_launch("powers_small");
}

// a * b + v: separately rounded in FP32 (the reference's float arithmetic,
// mirrored by the plain path), fused in FP64 (the precision paths).
__device__ __forceinline__ float madd(float a, float b, float v)
{
#pragma clang fp contract(off)
    const float pr = a * b;
    return v + pr;
}
__device__ __forceinline__ double madd(double a, double b, double v) { return fma(a, b, v); }

// ------------------------------------------------------------------ leaves
template <typename T>
struct Acc {
    using type = T;
};

// A work-group takes EB consecutive elements per thread group.  The sample
// rows its elements read form one contiguous run of U (element e reads rows
// e*sL + oL + q*rs, q < kq), staged once in LDS; each thread forms one row r
// (< n) of all EB leaves, so every Bbar value it loads is used EB times.
// Thread groups are 2^tpeLog >= n wide (<= 256); index math is shifts and
// products (no divisions).  Leaves past nReal are zero.  Rows of X outside
// [off, off + n) (the zero blocks of the ABM companion state) are written by
// each thread group for its own elements.
constexpr int kLeafEB = 8;

template <typename T, typename TU, typename TB>
__device__ __forceinline__ void d_leaves(const VBlock vb, int n, int m, int kq, int na, int off, long nReal,
                                                long nPad, int tpeLog, int groups,
                                                bool writeZeros, int zlo,
                                                const TB* __restrict__ BbT,
                                                const TU* __restrict__ U, long sL, long oL,
                                                long rs, T* __restrict__ X)
{
#pragma clang fp contract(off)
    using TA = decltype(TB(0) * TU(0));
    extern __shared__ unsigned char leaf_smem[];
    TA* Us = reinterpret_cast<TA*>(leaf_smem);       // staged rows x m
    const int tpe = 1 << tpeLog;
    const int grp = threadIdx.x >> tpeLog, lane = threadIdx.x & (tpe - 1);
    const long perBlock = (long)groups * kLeafEB;
    // the last sample row a real leaf reads
    const long lastRow = (nReal - 1) * sL + oL + (long)(kq - 1) * rs;
    const long es = sL * m;                           // staged values between elements
    for (long eb = (long)vb.x * perBlock; eb < nPad; eb += (long)vb.gx * perBlock) {
        const int cnt = (nPad - eb) < perBlock ? (int)(nPad - eb) : (int)perBlock;
        const long row0 = eb * sL + oL;
        long rows = (long)(cnt - 1) * sL + (long)(kq - 1) * rs + 1;
        if (rows > lastRow - row0 + 1) rows = lastRow - row0 + 1;
        const int vals = rows > 0 ? (int)(rows * m) : 0;
        const TU* src = U + row0 * m;
        // stage the block's sample rows (eight loads in flight per thread)
        constexpr int SU = 8;
        for (int t0 = 0; t0 < vals; t0 += 256 * SU) {
            TA v[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int t = t0 + threadIdx.x + u * 256;
                v[u] = t < vals ? (TA)src[t] : (TA)0;
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int t = t0 + threadIdx.x + u * 256;
                if (t < vals) Us[t] = v[u];
            }
        }
        const int le0 = grp * kLeafEB;
        const long e0 = eb + le0;
        if (grp < groups && le0 < cnt) {
            // zero rows [0, off) and [off + n, na) of the group's elements; when
            // the first level reads only rows from zlo on, rows [zlo, off) of
            // real elements and every zero row of padding elements
            const int ne = (cnt - le0) < kLeafEB ? cnt - le0 : kLeafEB;
            for (int u = 0; u < ne; ++u) {
                T* xe = X + (e0 + u) * na;
                const bool all = writeZeros || e0 + u >= nReal;
                for (int z = (all ? 0 : zlo) + lane; z < off; z += tpe) xe[z] = (T)0;
                if (all)
                    for (int z = off + n + lane; z < na; z += tpe) xe[z] = (T)0;
            }
        }
        __syncthreads();
        if (grp < groups && le0 < cnt) {
            for (int r = lane; r < n; r += tpe) {
                TA v[kLeafEB];
#pragma unroll
                for (int u = 0; u < kLeafEB; ++u) v[u] = 0;
                // Bbar columns c = q m + j in order, eight at a time (loads
                // issued together)
                for (int q = 0; q < kq; ++q) {
                    const TA* ub = Us + ((long)le0 * sL + (long)q * rs) * m;
                    const TB* bq = BbT + (long)q * m * n + r;
                    int j = 0;
                    for (; j + 8 <= m; j += 8) {
                        TA b[8];
#pragma unroll
                        for (int jj = 0; jj < 8; ++jj) b[jj] = (TA)bq[(long)(j + jj) * n];
#pragma unroll
                        for (int jj = 0; jj < 8; ++jj)
#pragma unroll
                            for (int u = 0; u < kLeafEB; ++u)
                                v[u] = madd(b[jj], ub[u * es + j + jj], v[u]);
                    }
                    for (; j < m; ++j) {
                        const TA b = (TA)bq[(long)j * n];
#pragma unroll
                        for (int u = 0; u < kLeafEB; ++u) v[u] = madd(b, ub[u * es + j], v[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < kLeafEB; ++u)
                    if (le0 + u < cnt)
                        X[(e0 + u) * na + off + r] = e0 + u < nReal ? (T)v[u] : (T)0;
            }
        }
        __syncthreads();
    }
}

template <typename T, typename TU, typename TB>
__global__ __launch_bounds__(256) void k_leaves(int n, int m, int kq, int na, int off, long nReal,
                                                long nPad, int tpeLog, int groups,
                                                bool writeZeros, int zlo,
                                                const TB* __restrict__ BbT,
                                                const TU* __restrict__ U, long sL, long oL,
                                                long rs, T* __restrict__ X)
{
    d_leaves<T, TU, TB>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, m, kq, na, off, nReal, nPad, tpeLog, groups, writeZeros, zlo, BbT, U, sL, oL, rs, X);
}

// FP64 leaves on the matrix cores (v_mfma_f64_16x16x4f64), as the transposed
// product V^T[e][r] = sum_c W[e][c] Bf[r][c] per tile of 16 leaves x 16 rows:
//   A (16 leaves x 4 columns): lane l holds W[e0 + l%16][c0 + l/16]
//   B (4 columns x 16 rows):   lane l holds Bf[r0 + l%16][c0 + l/16]
//   D: lane l, register v holds leaf e0 + 4v + l/16, row r0 + l%16
// The sample rows of the work-group's leaves (one contiguous run of U) are
// staged in LDS as FP64 with an odd row pitch (the 16 leaves of a fragment
// fall in distinct banks).  Without FAC, W is the leaf's samples themselves:
// column c = q m + j is input j of sample row e*sL + oL + q*rs (the lane's
// (q, j) advance by four columns per k-step without divisions) and Bf = Bbar.
// With FAC, Bbar = Bf (G (x) I_m) for a kq x kg matrix G (ABM: Bf = [B | AB]),
// and W[e][g m + j] = sum_q G[q][g] u_{e,q}[j] is formed in LDS first: kg m
// columns instead of kq m.  A wave takes one row tile of kLeafTiles leaf
// tiles; the work-group's waves cover the row tiles of G groups of leaf tiles.
// B fragments are loaded eight k-steps at a time.  Leaves past nReal are
// stored as zero.  Rows outside [off, off + n) are zero as in k_leaves.
constexpr int kLeafTiles = 4;
constexpr int kLeafMaxQ = 5, kLeafMaxG = 2;

template <typename T, typename TU, bool FAC>
__device__ __forceinline__ void d_leaves_mfma(const VBlock vb, int n, int m, int kq, int na, int off,
                                                     long nReal, long nPad, int groups,
                                                     int pitch, unsigned magic, int kg,
                                                     const LeafComb& G, bool writeZeros, int zlo,
                                                     const double* __restrict__ BbT,
                                                     const TU* __restrict__ U, long sL, long oL,
                                                     long rs, T* __restrict__ X)
{
    extern __shared__ unsigned char leaf_smem[];
    double* Us = reinterpret_cast<double*>(leaf_smem);   // staged rows x pitch
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, k4 = lane >> 4;
    const int km = FAC ? kg * m : kq * m, ks = (km + 3) / 4;
    const int rowTiles = (n + 15) / 16;
    const long perGroup = 16L * kLeafTiles, perBlock = perGroup * groups;
    const long lastRow = (nReal - 1) * sL + oL + (long)(kq - 1) * rs;
    const long stagedMax = (perBlock - 1) * sL + (long)(kq - 1) * rs + 1;
    const int pw = km | 1;                                  // W pitch (FAC)
    double* Ws = Us + stagedMax * pitch;                    // perBlock x pw (FAC)
    const int nz = na - n;
    // sample values of block eb staged into LDS (rows t / m by multiply-high)
    auto block_vals = [&](long eb) {
        const int cnt = (nPad - eb) < perBlock ? (int)(nPad - eb) : (int)perBlock;
        long rows = (long)(cnt - 1) * sL + (long)(kq - 1) * rs + 1;
        const long row0 = eb * sL + oL;
        if (rows > lastRow - row0 + 1) rows = lastRow - row0 + 1;
        return rows > 0 ? (int)(rows * m) : 0;
    };
    constexpr int SU = 8;
    // When one block's samples fit SU values per thread, the next block's are
    // loaded into registers while this block computes.
    const bool ahead = stagedMax * m <= 256L * SU;
    double v[SU];
    auto load_regs = [&](long eb) {
        const int vals = block_vals(eb);
        const TU* src = U + (eb * sL + oL) * m;
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int t = threadIdx.x + u * 256;
            v[u] = t < vals ? (double)src[t] : 0.0;
        }
    };
    auto store_regs = [&](int t0, int vals) {
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int t = t0 + threadIdx.x + u * 256;
            const int r = m == 1 ? t : (int)__umulhi((unsigned)t, magic);
            if (t < vals) Us[r * pitch + (t - r * m)] = v[u];
        }
    };
    const long stride = (long)vb.gx * perBlock;
    if (ahead && (long)vb.x * perBlock < nPad) load_regs((long)vb.x * perBlock);
    // When every wave has one row tile for the whole launch and K fits eight
    // k-steps, its B fragments stay in registers across blocks.
    const bool keepB = rowTiles * groups <= 4 && ks <= 8;
    double bk[8];
    if (keepB && wave < rowTiles * groups) {
        const int r0 = (wave % rowTiles) * 16;
        const int rr = (r0 + l16) < n ? r0 + l16 : n - 1;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int c = 4 * u + k4;
            bk[u] = c < km ? BbT[(long)c * n + rr] : 0.0;
        }
    }
    for (long eb = (long)vb.x * perBlock; eb < nPad; eb += stride) {
        const int cnt = (nPad - eb) < perBlock ? (int)(nPad - eb) : (int)perBlock;
        const int vals = block_vals(eb);
        if (ahead) {
            store_regs(0, vals);
        } else {
            // eight loads in flight per thread
            for (int t0 = 0; t0 < vals; t0 += 256 * SU) {
                const TU* src = U + (eb * sL + oL) * m;
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const int t = t0 + threadIdx.x + u * 256;
                    v[u] = t < vals ? (double)src[t] : 0.0;
                }
                store_regs(t0, vals);
            }
        }
        const int creal = nReal - eb < cnt ? (int)(nReal - eb > 0 ? nReal - eb : 0) : cnt;
        // zero rows [0, off) and [off + n, na); when the first level reads only
        // rows from zlo on, rows [zlo, off) of real elements and every zero row
        // of padding elements
        if (nz > 0) {
            // (leaf, row) pairs spread over the work-group: rows [zlo, off) of
            // the real leaves [partFrom, fullFrom), every zero row of [fullFrom, cnt)
            const int fullFrom = writeZeros ? 0 : creal;
            const int partFrom = (!writeZeros && zlo < off) ? 0 : fullFrom;
            const int w = off - zlo;
            for (int t = threadIdx.x; t < (fullFrom - partFrom) * w; t += 256) {
                const int q = t / w;
                X[(eb + partFrom + q) * na + zlo + (t - q * w)] = (T)0;
            }
            for (int t = threadIdx.x; t < (cnt - fullFrom) * nz; t += 256) {
                const int q = t / nz, z = t - q * nz;
                X[(eb + fullFrom + q) * na + (z < off ? z : z + n)] = (T)0;
            }
        }
        __syncthreads();
        if (ahead && eb + stride < nPad) load_regs(eb + stride);
        if (FAC) {
            // W[le][g m + j] = sum_q G[q][g] u[le*sL + q*rs][j], q in order
            for (int t = threadIdx.x; t < creal * m; t += 256) {
                const int le = m == 1 ? t : (int)__umulhi((unsigned)t, magic), j = t - le * m;
                const double* ur = Us + (long)le * sL * pitch + j;
                for (int g = 0; g < kg; ++g) {
                    double w = 0.0;
                    for (int q = 0; q < kq; ++q) w = fma(G.g[q][g], ur[(long)q * rs * pitch], w);
                    Ws[le * pw + g * m + j] = w;
                }
            }
            __syncthreads();
        }
        for (int item = wave; item < rowTiles * groups; item += 4) {
            const int rt = item % rowTiles, g = item / rowTiles;
            const int le0 = g * (int)perGroup;
            if (le0 >= cnt) continue;
            const int r0 = rt * 16;
            const int rr = (r0 + l16) < n ? r0 + l16 : n - 1;
            // LDS offset of this lane's leaf in each tile, clamped to a real
            // leaf of the block (products of the others are not stored)
            int ub[kLeafTiles];
#pragma unroll
            for (int t = 0; t < kLeafTiles; ++t) {
                int le = le0 + t * 16 + l16;
                le = le < creal - 1 ? le : (creal > 0 ? creal - 1 : 0);
                ub[t] = FAC ? le * pw : (int)(le * sL) * pitch;
            }
            doublex4 acc[kLeafTiles];
#pragma unroll
            for (int t = 0; t < kLeafTiles; ++t) acc[t] = doublex4{0.0, 0.0, 0.0, 0.0};
            int q = 0, j = k4;
            if (!FAC)
                while (j >= m) { j -= m; ++q; }
            for (int kb = 0; kb < ks; kb += 8) {
                double b[8];
                if (keepB) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) b[u] = bk[u];
                } else {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int c = 4 * (kb + u) + k4;
                        b[u] = c < km ? BbT[(long)c * n + rr] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (kb + u >= ks) break;
                    const int c = 4 * (kb + u) + k4;
                    const bool in = c < km;
                    const int cu = FAC ? c : (int)(q * rs) * pitch + j;
                    double a[kLeafTiles];
#pragma unroll
                    for (int t = 0; t < kLeafTiles; ++t) a[t] = in ? (FAC ? Ws : Us)[ub[t] + cu] : 0.0;
#pragma unroll
                    for (int t = 0; t < kLeafTiles; ++t)
                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[u], acc[t], 0, 0, 0);
                    if (!FAC) {
                        j += 4;
                        while (j >= m) { j -= m; ++q; }
                    }
                }
            }
            if (r0 + l16 >= n) continue;
#pragma unroll
            for (int t = 0; t < kLeafTiles; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int le = le0 + t * 16 + 4 * v + k4;
                    if (le < cnt)
                        X[(eb + le) * na + off + r0 + l16] = le < creal ? (T)acc[t][v] : (T)0;
                }
        }
        __syncthreads();
    }
}

template <typename T, typename TU, bool FAC>
__global__ __launch_bounds__(256, 3) void k_leaves_mfma(int n, int m, int kq, int na, int off,
                                                     long nReal, long nPad, int groups,
                                                     int pitch, unsigned magic, int kg,
                                                     LeafComb G, bool writeZeros, int zlo,
                                                     const double* __restrict__ BbT,
                                                     const TU* __restrict__ U, long sL, long oL,
                                                     long rs, T* __restrict__ X)
{
    d_leaves_mfma<T, TU, FAC>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, m, kq, na, off, nReal, nPad, groups, pitch, magic, kg, G, writeZeros, zlo, BbT, U, sL, oL, rs, X);
}

// LDS bytes of one work-group of k_leaves_mfma: the staged sample rows, and
// with kg > 0 the combined inputs W
static size_t leaves_mfma_lds(long perBlock, int m, int kq, int kg, long sL, long rs, int pitch)
{
    size_t d = (size_t)((perBlock - 1) * sL + (long)(kq - 1) * rs + 1) * pitch;
    if (kg > 0) d += (size_t)perBlock * ((kg * m) | 1);
    return d * sizeof(double);
}

template <typename T, typename TU>
static bool leaves_mfma(hipStream_t s, int n, int m, int kq, int na, int off, long nReal,
                        long nPad, const double* Bf, int kg, const LeafComb& G, const TU* U,
                        long sL, long oL, long rs, T* X, bool writeZeros)
{
    const int rowTiles = (n + 15) / 16;
    int groups = rowTiles >= 4 ? 1 : 4 / rowTiles;
    const int pitch = m | 1;
    while (groups > 1 &&
           leaves_mfma_lds(16L * kLeafTiles * groups, m, kq, kg, sL, rs, pitch) > 65536)
        groups >>= 1;
    const long perBlock = 16L * kLeafTiles * groups;
    const size_t lds = leaves_mfma_lds(perBlock, m, kq, kg, sL, rs, pitch);
    // the staged row index t / m (and the leaf index of W) is exact by
    // multiply-high for t, m < 2^16 (m = 1: the magic number 2^32 does not fit
    // 32 bits, and the kernel uses t itself)
    if (lds > 65536 || m >= 65536 || lds / sizeof(double) >= 65536 || perBlock * m >= 65536)
        return false;
    const unsigned magic = (unsigned)(((1ULL << 32) + m - 1) / m);
    // persistent work-groups (several blocks each, the next block's samples
    // loaded during the current one)
    const long blocks = std::min<long>((nPad + perBlock - 1) / perBlock, 2048);
    const dim3 grid((unsigned)std::max<long>(blocks, 1));
    if (kg > 0)
        hipLaunchKernelGGL((k_leaves_mfma<T, TU, true>), grid, dim3(256), lds, s, n, m, kq, na,
                           off, nReal, nPad, groups, pitch, magic, kg, G, writeZeros,
                           off / 32 * 32, Bf, U, sL, oL, rs, X);
    else
        hipLaunchKernelGGL((k_leaves_mfma<T, TU, false>), grid, dim3(256), lds, s, n, m, kq, na,
                           off, nReal, nPad, groups, pitch, magic, 0, G, writeZeros,
                           off / 32 * 32, Bf, U, sL, oL, rs, X);
    check_launch("leaves_mfma");
    return true;
}

void abm_leaf_factors(double h, LeafComb& G)
{
    // Bbar_q = G[q][0] B + G[q][1] AB, q = 0..4, as k_abm_combine forms it
    const AbmCoef c = abm_coef(h);
    const double g[5][2] = {{0.0, c.b3AB}, {c.b2B, c.b2AB}, {c.b1B, c.b1AB}, {c.b0B, c.b0AB},
                            {c.bm1, 0.0}};
    for (int q = 0; q < 5; ++q)
        for (int k = 0; k < 2; ++k) G.g[q][k] = g[q][k];
}

template <typename T, typename TU>
bool leaves_factored(hipStream_t s, int n, int m, int kq, int kg, int na, int off, long nReal,
                     long nPad, const double* BfT, const LeafComb& G, const TU* U, long sL,
                     long oL, long rs, T* X, bool writeZeros)
{
    if (m <= 0 || kq > kLeafMaxQ || kg < 1 || kg > kLeafMaxG || nReal < 1 || sL < 0 || oL < 0 ||
        rs < 0)
        return false;
    return leaves_mfma<T, TU>(s, n, m, kq, na, off, nReal, nPad, BfT, kg, G, U, sL, oL, rs, X,
                              writeZeros);
}

template bool leaves_factored<float, double>(hipStream_t, int, int, int, int, int, int, long,
                                             long, const double*, const LeafComb&,
                                             const double*, long, long, long, float*, bool);
template bool leaves_factored<float, float>(hipStream_t, int, int, int, int, int, int, long, long,
                                            const double*, const LeafComb&, const float*, long,
                                            long, long, float*, bool);
template bool leaves_factored<double, double>(hipStream_t, int, int, int, int, int, int, long,
                                              long, const double*, const LeafComb&,
                                              const double*, long, long, long, double*, bool);
template bool leaves_factored<double, float>(hipStream_t, int, int, int, int, int, int, long,
                                             long, const double*, const LeafComb&, const float*,
                                             long, long, long, double*, bool);

template <typename T, typename TU, typename TB>
void leaves(hipStream_t s, int n, int m, int kq, int na, int off, long nReal, long nPad,
            const TB* Bbar, const TU* U, long sL, long oL, long rs, T* X, bool writeZeros)
{
    using TA = decltype(TB(0) * TU(0));
    if (m <= 0) {   // no inputs: every leaf is zero
        hipLaunchKernelGGL((k_fill<T>), dim3(blocks_for(nPad * na, 256, 8192)), dim3(256), 0, s,
                           nPad * na, X);
        check_launch("leaves(zero)");
        return;
    }
    if (nReal < 1 || sL < 0 || oL < 0 || rs < 0)
        throw std::invalid_argument("leaves: bad sample layout");
    if constexpr (std::is_same<TB, double>::value) {   // FP64 products on the matrix cores
        if (leaves_mfma<T, TU>(s, n, m, kq, na, off, nReal, nPad, Bbar, 0, LeafComb{}, U, sL, oL,
                               rs, X, writeZeros))
            return;
    }
    int tpeLog = 0;
    while ((1 << tpeLog) < n && tpeLog < 8) ++tpeLog;
    // thread groups in use: as many as 256 threads hold, fewer when the staged
    // rows of their elements would pass 64 KiB of LDS
    int groups = 256 >> tpeLog;
    auto lds_for = [&](int g) {
        return (size_t)(((long)g * kLeafEB - 1) * sL + (long)(kq - 1) * rs + 1) * m * sizeof(TA);
    };
    while (groups > 1 && lds_for(groups) > 65536) groups >>= 1;
    if (lds_for(groups) > 65536) throw std::invalid_argument("leaves: too many inputs");
    const long perBlock = (long)groups * kLeafEB;
    const long blocks = std::min<long>((nPad + perBlock - 1) / perBlock, 1L << 16);
    hipLaunchKernelGGL((k_leaves<T, TU, TB>), dim3((unsigned)std::max<long>(blocks, 1)), dim3(256),
                       lds_for(groups), s, n, m, kq, na, off, nReal, nPad, tpeLog, groups,
                       writeZeros, off / 32 * 32, Bbar, U, sL, oL, rs, X);
    check_launch("leaves");
}

template void leaves<float, float, float>(hipStream_t, int, int, int, int, int, long, long,
                                          const float*, const float*, long, long, long, float*, bool);
template void leaves<float, double, double>(hipStream_t, int, int, int, int, int, long, long,
                                            const double*, const double*, long, long, long,
                                            float*, bool);
template void leaves<double, double, double>(hipStream_t, int, int, int, int, int, long, long,
                                             const double*, const double*, long, long, long,
                                             double*, bool);
template void leaves<double, float, double>(hipStream_t, int, int, int, int, int, long, long,
                                            const double*, const float*, long, long, long,
                                            double*, bool);

// ------------------------------------------------------------ scan levels
// One work-group owns TC whole column pairs (all na rows), so the down-sweep
// can update X[L] and X[R] in place.  The B operand (X[L] up, X[R] down) is
// staged in LDS with a padded pitch; the transposed level table is read
// row-coalesced (thread i reads PT[j][i]).  Thread t computes row
// i = t mod rowThreads (+ multiples of rowThreads) for the columns
// slot, slot + slots, ... (CPT of them).
constexpr int kLevelThreads = 256;

template <typename TX, typename TP, bool DOWN, bool DELTA, int kCPT>
__global__ __launch_bounds__(kLevelThreads) void k_level(int na, long d, long np, int slots,
                                                         int rowThreads, const TP* __restrict__ PT,
                                                         TX* __restrict__ X)
{
    extern __shared__ unsigned char smem_raw[];
    TP* Bs = reinterpret_cast<TP*>(smem_raw);
    const int TC = slots * kCPT, pitch = TC + 1;
    const long k0 = (long)blockIdx.x * TC;
    const int ncol = (np - k0) < (long)TC ? (int)(np - k0) : TC;
    for (int t = threadIdx.x; t < na * TC; t += kLevelThreads) {
        int c = t / na, i = t % na;
        TP v = 0;
        if (c < ncol) {
            long R = (k0 + c + 1) * 2 * d - 1;
            long src = DOWN ? R : R - d;
            v = (TP)X[src * na + i];
        }
        Bs[i * pitch + c] = v;
    }
    __syncthreads();
    const int t = threadIdx.x;
    const int row0 = t % rowThreads, slot = t / rowThreads;
    if (slot >= slots) return;
    for (int i = row0; i < na; i += rowThreads) {
        TP acc[kCPT];
#pragma unroll
        for (int q = 0; q < kCPT; ++q) acc[q] = 0;
        const TP* bcol = Bs + slot;
        int j = 0;
        for (; j + 4 <= na; j += 4) {
            TP p4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) p4[u] = PT[(long)(j + u) * na + i];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const TP* brow = bcol + (j + u) * pitch;
#pragma unroll
                for (int q = 0; q < kCPT; ++q) acc[q] = fma(p4[u], brow[q * slots], acc[q]);
            }
        }
        for (; j < na; ++j) {
            const TP p = PT[(long)j * na + i];
            const TP* brow = bcol + j * pitch;
#pragma unroll
            for (int q = 0; q < kCPT; ++q) acc[q] = fma(p, brow[q * slots], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < kCPT; ++q) {
            int c = slot + q * slots;
            if (c >= ncol) continue;
            long R = (k0 + c + 1) * 2 * d - 1, L = R - d;
            TP b = Bs[i * pitch + c];
            TP y = DELTA ? acc[q] + b : acc[q];
            if (!DOWN) {
                X[R * na + i] = (TX)((TP)X[R * na + i] + y);
            } else {
                TP xl = (TP)X[L * na + i];
                X[L * na + i] = (TX)b;
                X[R * na + i] = (TX)(y + xl);
            }
        }
    }
}

template <typename TX, typename TP, int CPT>
static void level_launch(hipStream_t s, bool down, bool delta, int na, long d, long np,
                         int slots, int rowThreads, const TP* PT, TX* X)
{
    int TC = slots * CPT;
    size_t lds = (size_t)na * (TC + 1) * sizeof(TP);
    unsigned grid = (unsigned)((np + TC - 1) / TC);
    dim3 g(grid), b(kLevelThreads);
    if (down) {
        if (delta)
            hipLaunchKernelGGL((k_level<TX, TP, true, true, CPT>), g, b, lds, s, na, d, np, slots,
                               rowThreads, PT, X);
        else
            hipLaunchKernelGGL((k_level<TX, TP, true, false, CPT>), g, b, lds, s, na, d, np, slots,
                               rowThreads, PT, X);
    } else {
        if (delta)
            hipLaunchKernelGGL((k_level<TX, TP, false, true, CPT>), g, b, lds, s, na, d, np, slots,
                               rowThreads, PT, X);
        else
            hipLaunchKernelGGL((k_level<TX, TP, false, false, CPT>), g, b, lds, s, na, d, np,
                               slots, rowThreads, PT, X);
    }
    check_launch("level");
}

template <typename TX, typename TP>
void level(hipStream_t s, bool down, bool delta, int na, long d, long np, const TP* PT, TX* X)
{
    if (np <= 0) return;
    int rowThreads = na >= kLevelThreads ? kLevelThreads : na;
    int slots = std::max(1, kLevelThreads / rowThreads);
    // largest columns-per-thread whose LDS stage fits in 64 KiB
    auto fits = [&](int cpt) { return (size_t)na * (slots * cpt + 1) * sizeof(TP) <= 65536; };
    if (fits(16))
        level_launch<TX, TP, 16>(s, down, delta, na, d, np, slots, rowThreads, PT, X);
    else if (fits(8))
        level_launch<TX, TP, 8>(s, down, delta, na, d, np, slots, rowThreads, PT, X);
    else if (fits(4))
        level_launch<TX, TP, 4>(s, down, delta, na, d, np, slots, rowThreads, PT, X);
    else
        level_launch<TX, TP, 2>(s, down, delta, na, d, np, slots, rowThreads, PT, X);
}

template void level<float, float>(hipStream_t, bool, bool, int, long, long, const float*, float*);
template void level<float, double>(hipStream_t, bool, bool, int, long, long, const double*,
                                   float*);
template void level<double, double>(hipStream_t, bool, bool, int, long, long, const double*,
                                    double*);

// ---------------------------------------------------- scan levels on MFMA
// FP32 level update on the matrix cores (v_mfma_f32_32x32x2_f32).  The product
// is formed transposed, Out[c][i] = sum_j Src[c][j] P[i][j] = (Src PT)[c][i],
// so that a 32-lane half of the accumulator holds 32 consecutive states i of
// one pair column c and the epilogue writes 128-byte runs of X.
//   A operand (32 pairs x 2 states):  the staged source columns, from LDS
//   B operand (2 states x 32 states): rows of the transposed table PT, from L2
// A work-group owns TC = WGC*NCT*32 whole pair columns (all na states), so the
// down-sweep stays in place.  Waves tile the states as WGI x NIT i-tiles and
// the pairs as WGC x NCT c-tiles; the source columns are staged KC states at a
// time with an odd pitch (conflict-free column reads).
typedef float floatx16 __attribute__((ext_vector_type(16)));

// PT is the zero-padded transposed table, [nrp][nap] with nap = 32*ceil(na/32)
// and nrp a multiple of KC, so fragment loads need no bounds checks; the next
// k-step's B fragments are loaded while the current MFMAs run.
// The lowest levels of the ABM scan have structure (LevelRows, all zero in
// general):
//   rbeg  first row of the X[L] source and X[R] accumulator (up) or the X[L]
//         addend (down) that is not zero; rows below are never read
//   jbeg  first source state the product needs (32 floor(rbeg / 32) up, 0 down)
//   ibase first row formed by the product (a multiple of 32); rows [clo, ibase)
//         are rows of P that are shifts, (P x)[i] = x[i + sh], and are copied;
//         below clo nothing is written (nothing reads those rows)
//   [zlo, zhi) rows of the up-sweep X[R] written as zero (the next level's
//         source rows between its jbeg and rbeg)
// Level 0 (P = M, leaves nonzero from 3n): rows below 3n of M are shifts by n.
// Level 1 (P = M^2, inputs nonzero from 2n): rows below 2n are shifts by 2n.
//
template <bool DOWN, bool DELTA, int NIT, int NCT, int WGI, int KC, bool SPLIT>
__global__ __launch_bounds__(256) void k_level_mfma(int na, int nap, int nrp, long d, long np,
                                                    LevelRows lr,
                                                    const float* __restrict__ PT,
                                                    const float* __restrict__ S,
                                                    float* __restrict__ X)
{
    constexpr int WGC = 4 / WGI;
    constexpr int TC = WGC * NCT * 32;
    constexpr int PITCH = KC + 1;
    constexpr int PER = TC * KC / 256;
    constexpr int IB = WGI * NIT * 32;
    __shared__ float As[TC * PITCH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wi = wave % WGI, wc = wave / WGI;
    const int l32 = lane & 31, hi = lane >> 5;
    const long k0 = (long)blockIdx.x * TC;
    const int jbeg = lr.jbeg, rbeg = lr.rbeg, ibase = lr.ibase;
    const int ib0 = ibase + (SPLIT ? (int)blockIdx.y * IB : 0);
    constexpr bool fromS = DOWN && SPLIT;
    const int ncol = (np - k0) < (long)TC ? (int)(np - k0) : TC;
    const bool full = ncol == TC;
    // Up-sweep: the accumulators start from the addend X[R], loaded before
    // the first K-slice so their latency hides behind it, and the epilogue
    // only stores; in the delta form they start at zero, so the small
    // products (P - I) X[L] are summed apart, and the epilogue adds
    // X[R] + X[L].  Down-sweep: they start at zero
    // and the epilogue loads X[R] and X[L] (a preload would raise the register
    // count past two waves per SIMD).  Columns past the tile are clamped
    // (loaded, never stored); level-0 leaf rows below rbeg are zero (one
    // branch per tile row).
    floatx16 acc[NCT][NIT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[ct][it][v] = 0.0f;
            const int i = ib0 + (wi + it * WGI) * 32 + l32;
            const int ic = i < na ? i : na - 1;
            if (!DOWN && !DELTA && i >= rbeg) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int cr = (wc + ct * WGC) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
                    const int crc = cr < ncol ? cr : ncol - 1;
                    const long R = (k0 + crc + 1) * 2 * d - 1;
                    acc[ct][it][v] = X[R * na + ic];
                }
            }
        }
    const float* pcol[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int col = ib0 + (wi + it * WGI) * 32 + l32;
        pcol[it] = PT + (col < nap ? col : nap - 1);
    }
    for (int j0 = jbeg; j0 < na; j0 += KC) {
        const int kc = (na - j0) < KC ? (na - j0) : KC;
        float v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int t = threadIdx.x + u * 256;
            const int c = t / KC, j = t % KC;
            // Clamped indices, so every load is in bounds and none is
            // predicated: columns past the tile repeat its last column (their
            // products are never stored), states past na repeat the last
            // state (they meet the zero rows of the padded table), and the
            // level-0 leaves hold zeros from jbeg on.
            const int cc = c < ncol ? c : ncol - 1, jj = j < kc ? j : kc - 1;
            const long R = (k0 + cc + 1) * 2 * d - 1;
            v[u] = fromS ? S[(k0 + cc) * na + j0 + jj] : X[(DOWN ? R : R - d) * na + j0 + jj];
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int t = threadIdx.x + u * 256;
            As[(t / KC) * PITCH + (t % KC)] = v[u];
        }
        __syncthreads();
        // k-steps of this slice: up to the 32-padded width (the padded table is
        // zero there and the staged columns are zero beyond na).  B fragments
        // are fetched KU k-steps at a time, one block ahead of the MFMAs that
        // use them; rows past the table are clamped (loaded, never used).
        constexpr int KU = 8;
        const int kend = (nap - j0) < KC ? (nap - j0) : KC;   // a multiple of 32
        float bc[KU][NIT], bn[KU][NIT];
#pragma unroll
        for (int u = 0; u < KU; ++u)
#pragma unroll
            for (int it = 0; it < NIT; ++it) bc[u][it] = pcol[it][(long)(j0 + 2 * u + hi) * nap];
        for (int kb = 0; kb < kend; kb += 2 * KU) {
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                int row = j0 + kb + 2 * KU + 2 * u + hi;
                row = row < nrp ? row : nrp - 1;
#pragma unroll
                for (int it = 0; it < NIT; ++it) bn[u][it] = pcol[it][(long)row * nap];
            }
            // keep the prefetch ahead of this block's MFMAs
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                const int kk = kb + 2 * u;
                float af[NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct)
                    af[ct] = As[((wc + ct * WGC) * 32 + l32) * PITCH + kk + hi];
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
                    for (int it = 0; it < NIT; ++it)
                        acc[ct][it] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[ct], bc[u][it],
                                                                           acc[ct][it], 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < KU; ++u)
#pragma unroll
                for (int it = 0; it < NIT; ++it) bc[u][it] = bn[u][it];
        }
        __syncthreads();
    }
    // shift rows [clo, ibase) and zero rows [zlo, zhi) of the tile's columns
    // (one row block does them); a wave per column, lanes over rows
    if ((lr.clo < ibase || lr.zlo < lr.zhi) && (!SPLIT || blockIdx.y == 0)) {
        for (int c = wave; c < ncol; c += 4) {
            const long R = (k0 + c + 1) * 2 * d - 1;
            for (int i = lr.clo + lane; i < ibase; i += 64) {
                if (!DOWN) {
                    X[R * na + i] = X[(R - d) * na + i + lr.sh];
                } else {
                    const float xs = fromS ? S[(k0 + c) * na + i + lr.sh] : X[R * na + i + lr.sh];
                    const float xr = fromS ? S[(k0 + c) * na + i] : X[R * na + i];
                    X[(R - d) * na + i] = xr;
                    X[R * na + i] = xs;
                }
            }
            if (!DOWN)
                for (int i = lr.zlo + lane; i < lr.zhi; i += 64) X[R * na + i] = 0.0f;
        }
        // the down-sweep copies read rows of X[R] that the epilogue overwrites
        if (DOWN && !fromS) __syncthreads();
    }
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int i = ib0 + (wi + it * WGI) * 32 + l32;
            if (i >= na) continue;
            if (!DOWN) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int cr = (wc + ct * WGC) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
                    if (!full && cr >= ncol) continue;
                    const long R = (k0 + cr + 1) * 2 * d - 1;
                    float y = acc[ct][it][v];
                    if (DELTA && i >= rbeg) y = X[R * na + i] + (X[(R - d) * na + i] + y);
                    X[R * na + i] = y;
                }
                continue;
            }
            // down: X[L] <- old X[R]; X[R] <- P old X[R] + old X[L]
            float xr[16], xl[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int cr = (wc + ct * WGC) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
                const int crc = cr < ncol ? cr : ncol - 1;
                xr[v] = fromS ? S[(k0 + crc) * na + i] : X[((k0 + crc + 1) * 2 * d - 1) * na + i];
                xl[v] = 0.0f;
            }
            if (i >= rbeg) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int cr = (wc + ct * WGC) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
                    const int crc = cr < ncol ? cr : ncol - 1;
                    xl[v] = X[((k0 + crc + 1) * 2 * d - 1 - d) * na + i];
                }
            }
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int cr = (wc + ct * WGC) * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
                if (!full && cr >= ncol) continue;
                const long R = (k0 + cr + 1) * 2 * d - 1;
                float y = acc[ct][it][v];
                if (DELTA) y = y + xr[v];
                X[(R - d) * na + i] = xr[v];
                X[R * na + i] = y + xl[v];
            }
        }
}

int mfma_table_rows(int na);

__global__ void k_gather_r(int na, long d, long np, const float* __restrict__ X,
                           float* __restrict__ S)
{
    const long total = np * na;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long c = e / na;
        const int j = (int)(e - c * na);
        S[e] = X[((c + 1) * 2 * d - 1) * na + j];
    }
}

template <int NIT, int NCT, int WGI, bool SPLIT>
static void mfma_launch_t(hipStream_t s, bool down, bool delta, int na, int nap, int nrp, long d,
                          long np, const LevelRows& lr, const float* PT, float* X,
                          const float* S, dim3 g)
{
    constexpr int TC = (4 / WGI) * NCT * 32;
    constexpr int KC = 8192 / TC;
    dim3 b(256);
    if (down) {
        if (delta)
            hipLaunchKernelGGL((k_level_mfma<true, true, NIT, NCT, WGI, KC, SPLIT>), g, b, 0, s, na,
                               nap, nrp, d, np, lr, PT, S, X);
        else
            hipLaunchKernelGGL((k_level_mfma<true, false, NIT, NCT, WGI, KC, SPLIT>), g, b, 0, s,
                               na, nap, nrp, d, np, lr, PT, S, X);
    } else {
        if (delta)
            hipLaunchKernelGGL((k_level_mfma<false, true, NIT, NCT, WGI, KC, SPLIT>), g, b, 0, s,
                               na, nap, nrp, d, np, lr, PT, S, X);
        else
            hipLaunchKernelGGL((k_level_mfma<false, false, NIT, NCT, WGI, KC, SPLIT>), g, b, 0, s,
                               na, nap, nrp, d, np, lr, PT, S, X);
    }
    check_launch("level_mfma");
}

template <int NIT, int NCT, int WGI>
static void mfma_launch(hipStream_t s, bool down, bool delta, int na, long d, long np,
                        const LevelRows& lr, const float* PT, float* X, float* S)
{
    constexpr int TC = (4 / WGI) * NCT * 32;
    constexpr int IB = WGI * NIT * 32;
    const int nap = (na + 31) / 32 * 32, nrp = mfma_table_rows(na);
    const unsigned gy = (unsigned)((na - lr.ibase + IB - 1) / IB);
    dim3 g((unsigned)((np + TC - 1) / TC), gy);
    if (gy == 1) {
        mfma_launch_t<NIT, NCT, WGI, false>(s, down, delta, na, nap, nrp, d, np, lr, PT, X,
                                            nullptr, g);
        return;
    }
    if (down) {
        if (!S) throw std::invalid_argument("level_mfma: split rows need the gather buffer");
        hipLaunchKernelGGL(k_gather_r, dim3(blocks_for(np * na, 256, 8192)), dim3(256), 0, s, na,
                           d, np, X, S);
        check_launch("gather_r");
    }
    mfma_launch_t<NIT, NCT, WGI, true>(s, down, delta, na, nap, nrp, d, np, lr, PT, X, S, g);
}

// Rows of the padded MFMA table: a multiple of every K-slice (<= 256).
int mfma_table_rows(int na) { return (na + 255) / 256 * 256; }

bool level_mfma_supported(int na) { return na >= 17 && na <= 1024; }

// Tile choice by scan size; the narrow variant (32 pairs per work-group)
// takes over when the wide one would leave the 256 CUs short of work.
void level_mfma(hipStream_t s, bool down, bool delta, int na, long d, long np, const float* PT,
                float* X, const LevelRows& lr, float* S, int tile)
{
    if (np <= 0) return;
    const bool shifts = lr.clo < lr.ibase, zeros = lr.zlo < lr.zhi;
    if (lr.jbeg % 32 != 0 || lr.jbeg < 0 || lr.jbeg >= na || lr.rbeg < 0 || lr.rbeg >= na ||
        lr.jbeg > lr.rbeg || (down && lr.jbeg != 0) || (delta && lr.rbeg != 0) ||
        lr.ibase % 32 != 0 || lr.ibase < 0 || lr.ibase >= na || lr.ibase > lr.rbeg ||
        lr.clo < 0 || lr.clo > lr.ibase || (shifts && (lr.sh <= 0 || lr.ibase + lr.sh > na)) ||
        (zeros && (down || lr.zlo < 0 || lr.zhi > lr.ibase || (shifts && lr.zhi > lr.clo))) ||
        (delta && (shifts || zeros)))
        throw std::invalid_argument("level_mfma: bad structured range");
    const int ibase = lr.ibase;
    const int nt = (na - ibase + 31) / 32;
    const bool narrow = tile == 1 || (tile == 0 && np < 2L * 256 * 64);
    if (nt <= 1) {
        if (narrow) mfma_launch<1, 1, 1>(s, down, delta, na, d, np, lr, PT, X, nullptr);
        else mfma_launch<1, 2, 1>(s, down, delta, na, d, np, lr, PT, X, nullptr);
    } else if (nt <= 2) {
        if (narrow) mfma_launch<1, 1, 2>(s, down, delta, na, d, np, lr, PT, X, nullptr);
        else mfma_launch<1, 2, 2>(s, down, delta, na, d, np, lr, PT, X, nullptr);
    } else if (nt <= 4) {
        if (narrow) mfma_launch<1, 1, 4>(s, down, delta, na, d, np, lr, PT, X, nullptr);
        else mfma_launch<1, 2, 4>(s, down, delta, na, d, np, lr, PT, X, nullptr);
    } else if (nt <= 8) {
        if (narrow) mfma_launch<2, 1, 4>(s, down, delta, na, d, np, lr, PT, X, nullptr);
        else mfma_launch<2, 2, 4>(s, down, delta, na, d, np, lr, PT, X, nullptr);
    } else {
        // above 256 rows: 256-row blocks side by side (blockIdx.y)
        mfma_launch<2, 2, 4>(s, down, delta, na, d, np, lr, PT, X, S);
    }
}

// ------------------------------------------------ ABM chunk recurrences
// The lowest lc levels of the ABM scan, replaced by a sequential recurrence
// per chunk of Lc = 2^lc elements.  The companion map shifts the state by one
// block per step, so a step needs only the n x 4n dense block row of M:
//   x_new[3n:4n) = x[3n:4n) + (M[3n:4n, :] - [0 0 0 I]) x + v_e[3n:4n),
//   x_new[0:3n)  = x[n:4n).
// The product is taken in delta form (the table is formed from the FP64 M,
// minus the identity, then rounded once), and the running sum of the newest
// block is compensated (Kahan): neither the rounding of M nor that of the
// state accumulates over the chunk.
// A work-group owns 16 chunks; their state vectors live in LDS as a ring of
// four blocks (the shift is a change of the ring's head), and each wave forms
// one 16-row tile of the new block on the matrix cores (v_mfma_f32_16x16x4f32):
//   A (16 chunks x 4 states): the ring, from LDS (lane l: chunk l%16)
//   B (4 states x 16 rows):   the delta table, held in registers for all k
//   D: lane l, register v holds chunk 4(l/16) + v, row l%16
// K-step ks = b NP/4 + q covers padded states b NP + q + g NP/4 (g = l/16),
// so a lane's A values of consecutive k-steps are adjacent in LDS.  Each lane
// also keeps its 4 entries of the newest block and their compensation terms
// in registers.  Blocks are padded to NP rows (the padded rows of the table
// and of the leaves are zero, so they stay zero).
//   Pass 1 (from zero): after Lc steps the ring is the chunk's aggregate, the
//   value the up-sweep leaves at X[(c+1)Lc - 1]; it is written there.
//   Pass 2 (after the down-sweep of the levels above): X[(c+1)Lc - 1] holds
//   x at the chunk's first element; the steps produce x for the others, and
//   the last block of each x_e is written to X[e][3n:4n) for the outputs.
//   Element e's leaf v_e is read from X[e] before x_e is written there.
constexpr int kChunkCols = 16;   // chunks per work-group

// LDS rows of the chunk kernels (one per chunk): lane group g = lane / 16 reads
// the KS entries [g KS, (g + 1) KS) of a row as 16-byte fragments, one row per
// lane of the group.  The LDS serves such a read 32 lanes at a time over 128
// banks (measured: the rows' old layout, groups at g KS with a pitch of 4 mod
// 128 words, had two bank conflicts per access unless KS = 64).  From KS = 12
// on, group g's entries start at chunk_base(g): 0, 64, 32, 96 (KS <= 32) or
// g 64 (KS <= 64), with a pitch of 4 mod 128 words, so that the 16 lanes of a
// group take distinct bank quads and the two groups served together take
// disjoint halves of the banks.  Narrower rows keep entries in order.
__host__ __device__ constexpr bool chunk_spread(int ks) { return ks >= 12 && ks <= 64; }
__host__ __device__ constexpr int chunk_pitch(int ks, int entries)
{
    return chunk_spread(ks) ? (ks <= 32 ? 132 : 260) : entries + 4;
}
__host__ __device__ constexpr int chunk_base(int ks, int g)
{
    return chunk_spread(ks) ? (ks <= 32 ? (g & 1) * 64 + (g >> 1) * 32 : g * 64) : g * ks;
}
// LDS offset of entry e of a row
__host__ __device__ constexpr int chunk_pos(int ks, int e)
{
    return chunk_base(ks, e / ks) + e % ks;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 4) void k_abm_chunks(int n, int lc, long nReal, long nChunks,
                                                       const double* __restrict__ P64,
                                                       float* __restrict__ X)
{
#pragma clang fp contract(off)
    constexpr int CH = kChunkCols;
    constexpr int KS = NP;                // k-steps of four states over the 4 NP padded states
    constexpr int PITCH = 4 * NP + 4;     // 4 mod 64 words; 16-byte aligned rows
    constexpr int NT = NP * 4;            // threads
    extern __shared__ float ring[];       // [CH][PITCH]
    const int na = 4 * n, Lc = 1 << lc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const int r = wave * 16 + l16;        // row of the new block this lane forms
    const bool rowOk = r < n;
    const long c0 = (long)blockIdx.x * CH;
    // B fragments (P64[k][i] = M[i][k]; padded state b NP + rr is the real
    // state b n + rr when rr < n)
    float bm[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int b = ks / (NP / 4), rr = ks % (NP / 4) + g * (NP / 4);
        double m = 0.0;
        if (rowOk && rr < n) m = P64[(long)(b * n + rr) * na + 3 * n + r] - (b == 3 && rr == r ? 1.0 : 0.0);
        bm[ks] = (float)m;
    }
    // chunks of the work-group that hold real elements
    long cReal = (nReal + Lc - 1) / Lc - c0;
    cReal = cReal < CH ? cReal : CH;
    if (cReal <= 0) return;
    // initial ring: zeros (pass 1) or the chunk's carry (pass 2), block b at slot b
    for (int t = threadIdx.x; t < CH * 4 * NP; t += NT) {
        const int c = t / (4 * NP), k = t % (4 * NP), b = k / NP, rr = k % NP;
        float v = 0.0f;
        if (PASS2 && c < cReal && rr < n) v = X[((c0 + c + 1) * Lc - 1) * na + b * n + rr];
        ring[c * PITCH + k] = v;
    }
    const int steps = PASS2 ? Lc - 1 : Lc;
    // this lane's chunks (clamped to real ones for the loads)
    int cl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = 4 * g + v;
        cl[v] = c < cReal ? c : (int)cReal - 1;
    }
    // leaves of the first step
    float vl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) vl[v] = rowOk ? X[((c0 + cl[v]) * Lc) * na + 3 * n + r] : 0.0f;
    __syncthreads();   // the ring is complete and the first leaves are in registers
    // the newest block and its compensation (the true value is xh - xc)
    float xh[4], xc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        xh[v] = rowOk ? ring[(4 * g + v) * PITCH + 3 * NP + r] : 0.0f;
        xc[v] = 0.0f;
    }
    // pass 2: the state of the chunk's first element goes out (over its leaf)
    if (PASS2 && rowOk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            const long e = (c0 + c) * Lc;
            if (c < cReal && e < nReal) X[e * na + 3 * n + r] = xh[v];
        }
    int head = 0;                         // slot of the oldest block
    for (int j = 0; j < steps; ++j) {
        // next step's leaves (element e + 1 of every chunk), issued before
        // the products; read before the state x_{e+1} is written over them
        float vn[4];
        const bool more = j + 1 < steps || !PASS2;
        const int jn = j + 1 < Lc ? j + 1 : Lc - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            vn[v] = (rowOk && more) ? X[((c0 + cl[v]) * Lc + jn) * na + 3 * n + r] : 0.0f;
        floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
        const float* arow = ring + l16 * PITCH + g * (NP / 4);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            // a block's A values first (16-byte LDS reads), then its products
            const floatx4* ab = reinterpret_cast<const floatx4*>(arow + ((head + b) & 3) * NP);
            floatx4 a[NP / 16];
#pragma unroll
            for (int q = 0; q < NP / 16; ++q) a[q] = ab[q];
#pragma unroll
            for (int q = 0; q < NP / 16; ++q) {
                const int ks = b * (NP / 4) + 4 * q;
                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][0], bm[ks], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][1], bm[ks + 1], acc1, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][2], bm[ks + 2], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][3], bm[ks + 3], acc1, 0, 0, 0);
            }
        }
        __syncthreads();                  // every wave has read the ring
        const int slot = head;            // the oldest block becomes the newest
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            // Kahan: xh + (delta + leaf), the lost low part kept in xc
            const float y = (acc0[v] + acc1[v] + vl[v]) - xc[v];
            const float sum = xh[v] + y;
            xc[v] = (sum - xh[v]) - y;
            xh[v] = sum;
            if (rowOk) ring[c * PITCH + slot * NP + r] = sum;
            if (PASS2 && rowOk) {
                const long e = (c0 + c) * Lc + j + 1;
                if (c < cReal && e < nReal) X[e * na + 3 * n + r] = sum - xc[v];
            }
            vl[v] = vn[v];
        }
        head = (head + 1) & 3;
        __syncthreads();
    }
    if (PASS2) return;
    // pass 1: the aggregate (ring, oldest block first; the newest block with
    // its compensation) to the chunk's last element
#pragma unroll
    for (int v = 0; v < 4; ++v)
        if (rowOk) ring[(4 * g + v) * PITCH + ((head + 3) & 3) * NP + r] = xh[v] - xc[v];
    __syncthreads();
    for (int t = threadIdx.x; t < CH * 4 * NP; t += NT) {
        const int c = t / (4 * NP), k = t % (4 * NP), b = k / NP, rr = k % NP;
        if (c < cReal && rr < n)
            X[((c0 + c + 1) * Lc - 1) * na + b * n + rr] = ring[c * PITCH + ((head + b) & 3) * NP + rr];
    }
}

// ABM chunk recurrences with an FP64 map, factored: the companion map's last
// block row is M[3n:4n, :] = [b0 A^2, a1 A + b1 A^2, a2 A + b2 A^2,
// I + a3 A + b3 A^2] (AdamsMoulton.cpp:258-269), so a step is
//   x_new = x3 + A w1 + A^2 w2 + v_e,  w1 = sum_b a_b x_b,  w2 = sum_b b_b x_b
// (x_b the state's blocks, oldest first; a_0 = 0): 2n products per row
// instead of 4n.  The lane that forms entry (chunk, row) of the new block
// keeps that entry's last four values in registers (the shift of the
// companion state), forms its w1 and w2 there and writes them to LDS, double
// buffered, so one barrier per step suffices; the newest value is a
// compensated running sum as in k_abm_chunks.  Tiles, lane layouts and the
// passes are those of k_abm_chunks, with A = [w1 w2] (16 chunks x 2 NP) and
// B = [A A^2]^T rows of the wave's tile, formed from the FP64 products and
// rounded once.
template <int NP, bool PASS2>
__device__ __forceinline__ void d_abm_chunks_fac(const VBlock vb, int n, int lc, long nReal,
                                                           const double* __restrict__ A64,
                                                           const double* __restrict__ A264,
                                                           AbmFactors f, float* __restrict__ X)
{
#pragma clang fp contract(off)
    constexpr int CH = kChunkCols;
    constexpr int KS = NP / 2;            // k-steps of four over the 2 NP padded [w1 w2]
    constexpr int PITCH = chunk_pitch(KS, 2 * NP);   // 16-byte aligned rows
    extern __shared__ float wbuf[];       // [2][CH][PITCH]
    const int na = 4 * n, Lc = 1 << lc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const int r = wave * 16 + l16;
    const bool rowOk = r < n;
    const long c0 = (long)vb.x * CH;
    // B fragments: lane group g takes the padded states [g KS, (g+1) KS) of
    // [w1 w2]; padded state b NP + rr (b = 0: A, 1: A^2) is real when rr < n
    float bm[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int kp = g * KS + ks, b = kp / NP, rr = kp % NP;
        double m = 0.0;
        if (rowOk && rr < n) m = (b == 0 ? A64 : A264)[(long)r * n + rr];
        bm[ks] = (float)m;
    }
    long cReal = (nReal + Lc - 1) / Lc - c0;
    cReal = cReal < CH ? cReal : CH;
    if (cReal <= 0) return;
    int cl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = 4 * g + v;
        cl[v] = c < cReal ? c : (int)cReal - 1;
    }
    // this lane's entries of the state's four blocks, oldest first, and the
    // compensation of the newest
    float hs[4][4], xc[4], xh[4], vl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const long last = ((c0 + cl[v] + 1) * Lc - 1) * na;
#pragma unroll
        for (int b = 0; b < 4; ++b) hs[v][b] = (PASS2 && rowOk) ? X[last + b * n + r] : 0.0f;
        xh[v] = hs[v][3];
        xc[v] = 0.0f;
        vl[v] = rowOk ? X[((c0 + cl[v]) * Lc) * na + 3 * n + r] : 0.0f;
    }
    auto put_w = [&](int buf) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            float w1 = 0.0f, w2 = 0.0f;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                w1 = w1 + f.a[b] * hs[v][b];
                w2 = w2 + f.b[b] * hs[v][b];
            }
            if (r < NP) {
                wbuf[(buf * CH + c) * PITCH + chunk_pos(KS, r)] = rowOk ? w1 : 0.0f;
                wbuf[(buf * CH + c) * PITCH + chunk_pos(KS, NP + r)] = rowOk ? w2 : 0.0f;
            }
        }
    };
    put_w(0);
    __syncthreads();   // w of the first step is in LDS, the first leaves in registers
    // pass 2: the state of the chunk's first element goes out (over its leaf)
    if (PASS2 && rowOk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            const long e = (c0 + c) * Lc;
            if (c < cReal && e < nReal) X[e * na + 3 * n + r] = hs[v][3];
        }
    const int steps = PASS2 ? Lc - 1 : Lc;
    int cur = 0;
    for (int j = 0; j < steps; ++j) {
        float vn[4];
        const bool more = j + 1 < steps || !PASS2;
        const int jn = j + 1 < Lc ? j + 1 : Lc - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            vn[v] = (rowOk && more) ? X[((c0 + cl[v]) * Lc + jn) * na + 3 * n + r] : 0.0f;
        floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
        const floatx4* ab =
            reinterpret_cast<const floatx4*>(wbuf + (cur * CH + l16) * PITCH + chunk_base(KS, g));
        floatx4 a[KS / 4];
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) a[q] = ab[q];
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][0], bm[4 * q], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][1], bm[4 * q + 1], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][2], bm[4 * q + 2], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][3], bm[4 * q + 3], acc1, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            // Kahan: x3 + (A w1 + A^2 w2 + leaf), the lost low part in xc
            const float y = (acc0[v] + acc1[v] + vl[v]) - xc[v];
            const float sum = xh[v] + y;
            xc[v] = (sum - xh[v]) - y;
            xh[v] = sum;
            const float xv = sum - xc[v];
#pragma unroll
            for (int b = 0; b < 3; ++b) hs[v][b] = hs[v][b + 1];
            hs[v][3] = xv;
            if (PASS2 && rowOk) {
                const long e = (c0 + c) * Lc + j + 1;
                if (c < cReal && e < nReal) X[e * na + 3 * n + r] = xv;
            }
            vl[v] = vn[v];
        }
        put_w(cur ^ 1);                   // the other buffer: no wave reads it this step
        cur ^= 1;
        __syncthreads();
    }
    if (PASS2) return;
    // pass 1: the aggregate (the four blocks) to the chunk's last element
    if (rowOk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            if (c < cReal)
#pragma unroll
                for (int b = 0; b < 4; ++b) X[((c0 + c + 1) * Lc - 1) * na + b * n + r] = hs[v][b];
        }
}

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 4) void k_abm_chunks_fac(int n, int lc, long nReal,
                                                           const double* __restrict__ A64,
                                                           const double* __restrict__ A264,
                                                           AbmFactors f, float* __restrict__ X)
{
    d_abm_chunks_fac<NP, PASS2>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, lc, nReal, A64, A264, f, X);
}

// The factored ABM step for 128 < n <= 256, as two dependent products:
//   x_new = x3 + A u + v_e,  u = w1 + A w2
// so a CU holds only A (n x n FP32, in registers: a wave's 32 rows of A over
// the whole padded state, split among its four lane groups) instead of
// [A A^2], which would fill the register file.  A work-group of NP / 32 waves
// takes 16 chunks; each wave forms two 16 x 16 tiles (rows wave*32 + 16 t),
// so one read of a 16 x 4 operand fragment from LDS feeds both.  Per step:
// w2 (written at the end of the previous step) -> A w2 -> u to LDS; barrier;
// A u -> compensated update, history shift, the next w1 (registers) and w2
// (LDS); barrier.  w2 and u have buffers of their own, so two barriers
// per step suffice.  Lane layouts and passes are those of k_abm_chunks_fac.
template <int NP, bool PASS2>
__device__ __forceinline__ void d_abm_chunks_two(const VBlock vb, int n, int lc, long nReal,
                                                           const double* __restrict__ A64,
                                                           AbmFactors f, float* __restrict__ X)
{
#pragma clang fp contract(off)
    constexpr int CH = kChunkCols;
    constexpr int KS = NP / 4;            // padded states per lane group
    constexpr int PITCH = chunk_pitch(KS, NP);   // 16-byte aligned rows
    static_assert(NP % 32 == 0 && KS % 4 == 0, "NP: a multiple of 32");
    extern __shared__ float sbuf[];       // w2: [CH][PITCH], then u: [CH][PITCH]
    float* const w2buf = sbuf;
    float* const ubuf = sbuf + CH * PITCH;
    const int na = 4 * n, Lc = 1 << lc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    int rr[2];
    bool ok[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        rr[t] = wave * 32 + 16 * t + l16;
        ok[t] = rr[t] < n;
    }
    const long c0 = (long)vb.x * CH;
    float bm[2][KS];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = g * KS + ks;
            bm[t][ks] = (ok[t] && k < n) ? (float)A64[(long)rr[t] * n + k] : 0.0f;
        }
    long cReal = (nReal + Lc - 1) / Lc - c0;
    cReal = cReal < CH ? cReal : CH;
    if (cReal <= 0) return;
    int cl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = 4 * g + v;
        cl[v] = c < cReal ? c : (int)cReal - 1;
    }
    float hs[2][4][4], xc[2][4], xh[2][4], vl[2][4], w1[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long last = ((c0 + cl[v] + 1) * Lc - 1) * na;
#pragma unroll
            for (int b = 0; b < 4; ++b)
                hs[t][v][b] = (PASS2 && ok[t]) ? X[last + b * n + rr[t]] : 0.0f;
            xh[t][v] = hs[t][v][3];
            xc[t][v] = 0.0f;
            vl[t][v] = ok[t] ? X[((c0 + cl[v]) * Lc) * na + 3 * n + rr[t]] : 0.0f;
        }
    // w1 stays in this lane's registers; w2 goes to LDS for the first product
    auto put_w = [&]() {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float a = 0.0f, b2 = 0.0f;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    a = a + f.a[b] * hs[t][v][b];
                    b2 = b2 + f.b[b] * hs[t][v][b];
                }
                w1[t][v] = a;
                w2buf[(4 * g + v) * PITCH + chunk_pos(KS, rr[t])] = ok[t] ? b2 : 0.0f;
            }
    };
    // 16 chunks x NP padded states of `buf` times this wave's rows of A
    auto product = [&](const float* buf, floatx4 (&acc)[2]) {
        acc[0] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        acc[1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        const floatx4* ab =
            reinterpret_cast<const floatx4*>(buf + l16 * PITCH + chunk_base(KS, g));
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
            const floatx4 a = ab[q];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bm[0][4 * q + i], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bm[1][4 * q + i], acc[1], 0, 0, 0);
            }
        }
    };
    put_w();
    __syncthreads();
    if (PASS2)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 4 * g + v;
                const long e = (c0 + c) * Lc;
                if (ok[t] && c < cReal && e < nReal) X[e * na + 3 * n + rr[t]] = hs[t][v][3];
            }
    const int steps = PASS2 ? Lc - 1 : Lc;
    for (int j = 0; j < steps; ++j) {
        float vn[2][4];
        const bool more = j + 1 < steps || !PASS2;
        const int jn = j + 1 < Lc ? j + 1 : Lc - 1;
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                vn[t][v] = (ok[t] && more) ? X[((c0 + cl[v]) * Lc + jn) * na + 3 * n + rr[t]] : 0.0f;
        floatx4 acc[2];
        product(w2buf, acc);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                ubuf[(4 * g + v) * PITCH + chunk_pos(KS, rr[t])] =
                    ok[t] ? w1[t][v] + acc[t][v] : 0.0f;
        __syncthreads();
        product(ubuf, acc);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 4 * g + v;
                const float y = (acc[t][v] + vl[t][v]) - xc[t][v];
                const float sum = xh[t][v] + y;
                xc[t][v] = (sum - xh[t][v]) - y;
                xh[t][v] = sum;
                const float xv = sum - xc[t][v];
#pragma unroll
                for (int b = 0; b < 3; ++b) hs[t][v][b] = hs[t][v][b + 1];
                hs[t][v][3] = xv;
                if (PASS2 && ok[t]) {
                    const long e = (c0 + c) * Lc + j + 1;
                    if (c < cReal && e < nReal) X[e * na + 3 * n + rr[t]] = xv;
                }
                vl[t][v] = vn[t][v];
            }
        put_w();                          // every wave has read w2 this step
        __syncthreads();
    }
    if (PASS2) return;
#pragma unroll
    for (int t = 0; t < 2; ++t)
        if (ok[t])
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 4 * g + v;
                if (c < cReal)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        X[((c0 + c + 1) * Lc - 1) * na + b * n + rr[t]] = hs[t][v][b];
            }
}

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 2) void k_abm_chunks_two(int n, int lc, long nReal,
                                                           const double* __restrict__ A64,
                                                           AbmFactors f, float* __restrict__ X)
{
    d_abm_chunks_two<NP, PASS2>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, lc, nReal, A64, f, X);
}

bool abm_chunks_factored_supported(int n) { return n >= 1 && n <= 256; }

AbmFactors abm_chunk_factors(double h)
{
    const AbmCoef c = abm_coef(h);
    // blocks oldest first: x_{k-3} (A^2 only), x_{k-2}, x_{k-1}, x_k (+ I)
    return AbmFactors{{0.0f, (float)c.a2A, (float)c.a1A, (float)c.a0A},
                      {(float)c.a3A2, (float)c.a2A2, (float)c.a1A2, (float)c.a0A2}};
}

void abm_chunks_factored(hipStream_t s, bool pass2, int n, int lc, long nReal, long nPad,
                         double h, const double* A64, const double* A264, float* X)
{
    if (!abm_chunks_factored_supported(n) || lc < 1 || (nPad >> lc) < 1 || nReal < 1 ||
        nReal > nPad)
        throw std::invalid_argument("abm_chunks_factored: bad shape");
    const AbmFactors f = abm_chunk_factors(h);
    const long nChunks = nPad >> lc;
    const dim3 grid((unsigned)((nChunks + kChunkCols - 1) / kChunkCols));
    auto launch = [&](auto np) {
        constexpr int NP = decltype(np)::value;
        const size_t lds = (size_t)2 * kChunkCols * chunk_pitch(NP / 2, 2 * NP) * sizeof(float);
        if (pass2)
            hipLaunchKernelGGL((k_abm_chunks_fac<NP, true>), grid, dim3(NP * 4), lds, s, n, lc,
                               nReal, A64, A264, f, X);
        else
            hipLaunchKernelGGL((k_abm_chunks_fac<NP, false>), grid, dim3(NP * 4), lds, s, n, lc,
                               nReal, A64, A264, f, X);
    };
    auto launch_two = [&](auto np) {
        constexpr int NP = decltype(np)::value;
        const size_t lds = (size_t)2 * kChunkCols * chunk_pitch(NP / 4, NP) * sizeof(float);
        if (pass2)
            hipLaunchKernelGGL((k_abm_chunks_two<NP, true>), grid, dim3(NP * 2), lds, s, n, lc,
                               nReal, A64, f, X);
        else
            hipLaunchKernelGGL((k_abm_chunks_two<NP, false>), grid, dim3(NP * 2), lds, s, n, lc,
                               nReal, A64, f, X);
    };
    if (n <= 16)
        launch(std::integral_constant<int, 16>());
    else if (n <= 32)
        launch(std::integral_constant<int, 32>());
    else if (n <= 48)
        launch(std::integral_constant<int, 48>());
    else if (n <= 64)
        launch(std::integral_constant<int, 64>());
    else if (n <= 96)
        launch(std::integral_constant<int, 96>());
    else if (n <= 128)
        launch(std::integral_constant<int, 128>());
    else if (n <= 160)
        launch_two(std::integral_constant<int, 160>());
    else if (n <= 192)
        launch_two(std::integral_constant<int, 192>());
    else if (n <= 224)
        launch_two(std::integral_constant<int, 224>());
    else
        launch_two(std::integral_constant<int, 256>());
    check_launch("abm_chunks_factored");
}

// Chunk recurrences of a dense map (RK4's M, or the ABM power P_l over chunk
// aggregates): x_new = x + (P - I) x + v_e over the whole state, for elements
// at X rows pos(e) = e S + S - 1 (S = 1: every element; S = 2^l: the values
// the tree holds after level l - 1).  As in k_abm_chunks: a work-group owns 16
// chunks with their states in LDS, a wave forms one 16-row tile of the new
// state on v_mfma_f32_16x16x4f32 with its rows of P - I (formed from the FP64
// table, then rounded once) in registers, and every state entry is a
// compensated running sum.  Lane group g = l/16 takes the states
// [g NP/4, (g+1) NP/4) of each k-step, so its A values are adjacent in LDS.
//   Pass 1 (from zero) writes each chunk's aggregate to its last element's
//   row; pass 2 starts from the value there (the state at the chunk's first
//   element) and writes the state of every element of the chunk to its row,
//   each after reading that element's leaf.
template <int NP, bool PASS2>
__device__ __forceinline__ void d_chunks_dense(const VBlock vb, int na, int lc, long S, long nReal,
                                                         const double* __restrict__ P64,
                                                         float* __restrict__ X)
{
#pragma clang fp contract(off)
    constexpr int CH = kChunkCols;
    constexpr int KS = NP / 4;            // k-steps of four states
    constexpr int PITCH = chunk_pitch(KS, NP);   // 16-byte aligned rows
    constexpr int NT = NP * 4;
    extern __shared__ float xsd[];        // [CH][PITCH]
    const int Lc = 1 << lc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const int r = wave * 16 + l16;
    const bool rowOk = r < na;
    const long c0 = (long)vb.x * CH;
    auto pos = [&](long e) { return e * S + S - 1; };
    float bm[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int st = g * KS + ks;
        double m = 0.0;
        if (rowOk && st < na) m = P64[(long)st * na + r] - (st == r ? 1.0 : 0.0);
        bm[ks] = (float)m;
    }
    long cReal = (nReal + Lc - 1) / Lc - c0;
    cReal = cReal < CH ? cReal : CH;
    if (cReal <= 0) return;
    int cl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = 4 * g + v;
        cl[v] = c < cReal ? c : (int)cReal - 1;
    }
    // initial state: zeros or the chunk's carry; first leaves
    float xh[4], xc[4], vl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const long last = pos((c0 + cl[v] + 1) * Lc - 1);
        xh[v] = (PASS2 && rowOk) ? X[last * na + r] : 0.0f;
        xc[v] = 0.0f;
        vl[v] = rowOk ? X[pos((c0 + cl[v]) * Lc) * na + r] : 0.0f;
    }
    for (int t = threadIdx.x; t < CH * PITCH; t += NT) xsd[t] = 0.0f;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; ++v)
        if (rowOk) xsd[(4 * g + v) * PITCH + chunk_pos(KS, r)] = xh[v];
    // pass 2: the state of the chunk's first element goes out (over its leaf)
    if (PASS2 && rowOk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            const long e = (c0 + c) * Lc;
            if (c < cReal && e < nReal) X[pos(e) * na + r] = xh[v];
        }
    __syncthreads();
    const int steps = PASS2 ? Lc - 1 : Lc;
    for (int j = 0; j < steps; ++j) {
        float vn[4];
        const bool more = j + 1 < steps || !PASS2;
        const int jn = j + 1 < Lc ? j + 1 : Lc - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            vn[v] = (rowOk && more) ? X[pos((c0 + cl[v]) * Lc + jn) * na + r] : 0.0f;
        floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
        const floatx4* ab =
            reinterpret_cast<const floatx4*>(xsd + l16 * PITCH + chunk_base(KS, g));
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
            const floatx4 a = ab[q];
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], bm[4 * q], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], bm[4 * q + 1], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], bm[4 * q + 2], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], bm[4 * q + 3], acc1, 0, 0, 0);
        }
        __syncthreads();                  // every wave has read the state
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * g + v;
            const float y = (acc0[v] + acc1[v] + vl[v]) - xc[v];
            const float sum = xh[v] + y;
            xc[v] = (sum - xh[v]) - y;
            xh[v] = sum;
            if (rowOk) xsd[c * PITCH + chunk_pos(KS, r)] = sum;
            if (PASS2 && rowOk) {
                const long e = (c0 + c) * Lc + j + 1;
                if (c < cReal && e < nReal) X[pos(e) * na + r] = sum - xc[v];
            }
            vl[v] = vn[v];
        }
        __syncthreads();
    }
    if (PASS2) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = 4 * g + v;
        if (rowOk && c < cReal) X[pos((c0 + c + 1) * Lc - 1) * na + r] = xh[v] - xc[v];
    }
}

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 4) void k_chunks_dense(int na, int lc, long S, long nReal,
                                                         const double* __restrict__ P64,
                                                         float* __restrict__ X)
{
    d_chunks_dense<NP, PASS2>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, lc, S, nReal, P64, X);
}

bool chunks_dense_supported(int na) { return na >= 1 && na <= 256; }

void chunks_dense(hipStream_t s, bool pass2, int na, int lc, long S, long nReal, long nElem,
                  const double* P64, float* X)
{
    if (!chunks_dense_supported(na) || lc < 1 || S < 1 || (nElem >> lc) < 1 || nReal < 1 ||
        nReal > nElem)
        throw std::invalid_argument("chunks_dense: bad shape");
    const long nChunks = nElem >> lc;
    const dim3 grid((unsigned)((nChunks + kChunkCols - 1) / kChunkCols));
    auto launch = [&](auto np) {
        constexpr int NP = decltype(np)::value;
        const size_t lds = (size_t)kChunkCols * chunk_pitch(NP / 4, NP) * sizeof(float);
        if (pass2)
            hipLaunchKernelGGL((k_chunks_dense<NP, true>), grid, dim3(NP * 4), lds, s, na, lc, S,
                               nReal, P64, X);
        else
            hipLaunchKernelGGL((k_chunks_dense<NP, false>), grid, dim3(NP * 4), lds, s, na, lc, S,
                               nReal, P64, X);
    };
    if (na <= 16)
        launch(std::integral_constant<int, 16>());
    else if (na <= 32)
        launch(std::integral_constant<int, 32>());
    else if (na <= 48)
        launch(std::integral_constant<int, 48>());
    else if (na <= 64)
        launch(std::integral_constant<int, 64>());
    else if (na <= 96)
        launch(std::integral_constant<int, 96>());
    else if (na <= 128)
        launch(std::integral_constant<int, 128>());
    else if (na <= 160)
        launch(std::integral_constant<int, 160>());
    else if (na <= 192)
        launch(std::integral_constant<int, 192>());
    else if (na <= 224)
        launch(std::integral_constant<int, 224>());
    else
        launch(std::integral_constant<int, 256>());
    check_launch("chunks_dense");
}

bool abm_chunks_supported(int n) { return n >= 1 && n <= 64; }

void abm_chunks(hipStream_t s, bool pass2, int n, int lc, long nReal, long nPad,
                const double* P64, float* X)
{
    if (!abm_chunks_supported(n) || lc < 1 || (nPad >> lc) < 1 || nReal < 1 || nReal > nPad)
        throw std::invalid_argument("abm_chunks: bad shape");
    const long nChunks = nPad >> lc;
    const dim3 grid((unsigned)((nChunks + kChunkCols - 1) / kChunkCols));
    auto launch = [&](auto np) {
        constexpr int NP = decltype(np)::value;
        const size_t lds = (size_t)kChunkCols * (4 * NP + 4) * sizeof(float);
        if (pass2)
            hipLaunchKernelGGL((k_abm_chunks<NP, true>), grid, dim3(NP * 4), lds, s, n, lc, nReal,
                               nChunks, P64, X);
        else
            hipLaunchKernelGGL((k_abm_chunks<NP, false>), grid, dim3(NP * 4), lds, s, n, lc, nReal,
                               nChunks, P64, X);
    };
    if (n <= 16)
        launch(std::integral_constant<int, 16>());
    else if (n <= 32)
        launch(std::integral_constant<int, 32>());
    else
        launch(std::integral_constant<int, 64>());
    check_launch("abm_chunks");
}

// ------------------------------------------- levels with few pairs (K split)
// A work-group computes one tile of pairs x rows (32 x 32 on FP32 MFMA,
// 16 x 16 on FP64 MFMA); its four waves split the K range (the source states)
// and their partial tiles meet in LDS.  The many small tiles spread a level
// with few pairs over many CUs, where one work-group per column block would
// stream the whole table through one CU.  Sources: X[L] (up) or S, a copy of
// the old X[R] (down; other tiles write X[R] rows meanwhile).  Tables: the
// padded transposed FP32 table [nrp][nap] or the transposed FP64 power
// [na][na].  Up: X[R] += P X[L]; down: X[L] = old X[R], X[R] = P old X[R] +
// X[L]; with DELTA (tables P - I) the sum is the small part alone and the
// epilogue adds it to X[R] + X[L] (up) or X[L] + old X[R] (down).  Loads are
// clamped into range and issued eight k-steps ahead.
// SMALL (FP32 only): 16 x 16 tiles on v_mfma_f32_16x16x4f32, four times the
// work-groups of the 32 x 32 tiles, for levels whose 32 x 32 grid would leave
// most CUs idle.
template <typename TT, bool SMALL>
struct KsTile;
template <>
struct KsTile<float, false> {
    static constexpr int T = 32, KS = 2, NACC = 16;
    using acc_t = floatx16;
};
template <>
struct KsTile<float, true> {
    static constexpr int T = 16, KS = 4, NACC = 4;
    using acc_t = floatx4;
};
template <>
struct KsTile<double, false> {
    static constexpr int T = 16, KS = 4, NACC = 4;
    using acc_t = doublex4;
};

template <typename TT, typename TX, bool DOWN, bool DELTA, bool SMALL>
__device__ __forceinline__ void d_level_ks(const VBlock vb, int na, int ldp, long d, long np,
                                                  const TT* __restrict__ PT,
                                                  const TX* __restrict__ S, TX* __restrict__ X,
                                                  TX* __restrict__ Sn, long npn)
{
    using K = KsTile<TT, SMALL>;
    constexpr int T = K::T, KS = K::KS, NACC = K::NACC, U = 8;
    // fragment and result layouts: 32 x 32 x 2 FP32 (T = 32), or the 16 x 16 x 4
    // shapes (T = 16), whose A and B layouts agree and whose D layouts differ
    constexpr bool F32BIG = sizeof(TT) == 4 && !SMALL;
    __shared__ TT part[3][NACC][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long c0 = (long)vb.x * T;
    const int i0 = vb.y * T;
    // fragment coordinates: FP32 A(c = l%32, k = l/32), B(k = l/32, i = l%32);
    // FP64 A(c = l%16, k = l/16), B(k = l/16, i = l%16)
    const int lc = lane % T, lk = lane / T;
    const long colA = (c0 + lc) < np ? c0 + lc : np - 1;
    const TX* src = DOWN ? S + colA * na : X + ((colA + 1) * 2 * d - 1 - d) * na;
    const int ib = i0 + lc;
    const int ibc = (sizeof(TT) == 4 || ib < na) ? ib : na - 1;   // FP32 table is padded
    const int ksteps = (na + KS - 1) / KS, per = (ksteps + 3) / 4;
    const int kb = wave * per, ke = (kb + per) < ksteps ? kb + per : ksteps;
    typename K::acc_t acc;
#pragma unroll
    for (int v = 0; v < NACC; ++v) acc[v] = 0;
    for (int k = kb; k < ke; k += U) {
        TT a[U], b[U];
        TX araw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // clamped, unpredicated loads, all issued before the MFMAs; a
            // k-step past the range or a state past na contributes a zero A
            // value (the loaded values are finite)
            const int j = (k + u) * KS + lk;
            const int jc = j < na ? j : na - 1;
            araw[u] = src[jc];
            b[u] = PT[(long)jc * ldp + ibc];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = (k + u) * KS + lk;
            a[u] = (k + u < ke && j < na) ? (TT)araw[u] : (TT)0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (F32BIG)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[u], acc, 0, 0, 0);
            else if constexpr (sizeof(TT) == 4)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], acc, 0, 0, 0);
            else
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int v = 0; v < NACC; ++v) part[wave - 1][v][lane] = acc[v];
    }
    __syncthreads();
    if (wave != 0) return;
    // D: FP32 32x32 (c = 8(v/4) + 4(l/32) + v%4, i = l%32); FP32 16x16
    // (c = 4(l/16) + v, i = l%16); FP64 (c = 4v + l/16, i = l%16).
    // All loads first (clamped addresses), then the guarded stores.
    const int i = i0 + lc, icl = i < na ? i : na - 1;
    TX xo[NACC], xs[NACC];
#pragma unroll
    for (int v = 0; v < NACC; ++v) {
        const int c = F32BIG ? 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)
                             : sizeof(TT) == 4 ? 4 * (lane >> 4) + v : 4 * v + (lane >> 4);
        const long col = (c0 + c) < np ? c0 + c : np - 1;
        const long R = (col + 1) * 2 * d - 1;
        xo[v] = X[(DOWN ? R - d : R) * na + icl];
        xs[v] = DOWN ? S[col * na + icl] : DELTA ? X[(R - d) * na + icl] : (TX)0;
    }
#pragma unroll
    for (int v = 0; v < NACC; ++v) {
        const TT sum = acc[v] + part[0][v][lane] + part[1][v][lane] + part[2][v][lane];
        const int c = F32BIG ? 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)
                             : sizeof(TT) == 4 ? 4 * (lane >> 4) + v : 4 * v + (lane >> 4);
        const long col = c0 + c;
        if (col >= np || i >= na) continue;
        const long R = (col + 1) * 2 * d - 1;
        if (!DOWN) {
            X[R * na + i] = DELTA ? (TX)((TT)xo[v] + ((TT)xs[v] + sum)) : (TX)((TT)xo[v] + sum);
        } else {
            const TX y = DELTA ? (TX)((TT)xo[v] + ((TT)xs[v] + sum)) : (TX)(sum + (TT)xo[v]);
            X[(R - d) * na + i] = xs[v];
            X[R * na + i] = y;
            // the next (lower) level's copy of its pairs' X[R]: its pairs 2 col
            // and 2 col + 1 end at this pair's L and R
            if (Sn) {
                if (2 * col < npn) Sn[(2 * col) * na + i] = xs[v];
                if (2 * col + 1 < npn) Sn[(2 * col + 1) * na + i] = y;
            }
        }
    }
}

// FP32 levels with many pairs as 64 x 64 tiles (64 pairs x 64 table rows per
// work-group; four waves, each a 32 x 32 quadrant on v_mfma_f32_32x32x2f32),
// K (the source states) in slices of 32 staged in LDS and double buffered:
// the pairs' source rows and the table's rows arrive in coalesced runs, where
// the K-split tiles gather every source value with its own strided load.
// Sources, tables and the epilogue are those of d_level_ks (up: X[R] += P
// X[L]; down from S, the copy of the old X[R]; DELTA tables P - I; Sn, the
// next lower level's copy).  Grid: x = pair tile, y = row tile.
constexpr int kLvG = 64, kLvKB = 32;

template <bool DOWN, bool DELTA>
__device__ __forceinline__ void d_level_gemm(const VBlock vb, int na, int ldp, long d, long np,
                                             const float* __restrict__ PT,
                                             const float* __restrict__ S, float* __restrict__ X,
                                             float* __restrict__ Sn, long npn)
{
    __shared__ float As[2][kLvG][kLvKB + 1];
    __shared__ float Bs[2][kLvKB][kLvG + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const long c0 = (long)vb.x * kLvG;
    const int i0 = (int)vb.y * kLvG;
    const int ar = tid >> 2, ak = (tid & 3) * 8;     // A slice: pair, first of eight states
    const int bk = tid >> 3, bj = (tid & 7) * 8;     // B slice: state, first of eight rows
    const long colA = (c0 + ar) < np ? c0 + ar : np - 1;
    const float* src = DOWN ? S + colA * na : X + ((colA + 1) * 2 * d - 1 - d) * na;
    float ra[8], rb[8];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + ak + q;
            ra[q] = k < na ? src[k] : 0.0f;
            const int kb = k0 + bk, i = i0 + bj + q;
            rb[q] = (kb < na && i < ldp) ? PT[(long)kb * ldp + i] : 0.0f;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            As[buf][ar][ak + q] = ra[q];
            Bs[buf][bk][bj + q] = rb[q];
        }
    };
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    const int l32 = lane & 31, h = lane >> 5;
    for (int k0 = 0; k0 < na; k0 += kLvKB) {
        const bool more = k0 + kLvKB < na;
        if (more) load(k0 + kLvKB);
#pragma unroll
        for (int ks = 0; ks < kLvKB; ks += 2) {
            // A(c = l%32, k = l/32), B(k = l/32, i = l%32)
            const float a = As[buf][wr * 32 + l32][ks + h];
            const float b = Bs[buf][ks + h][wc * 32 + l32];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // D: c = 8 (v / 4) + 4 (l / 32) + v % 4, i = l % 32 within the quadrant;
    // all loads first (clamped addresses), then the guarded stores
    const int i = i0 + wc * 32 + l32, icl = i < na ? i : na - 1;
    float xo[16], xs[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const long c = c0 + wr * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
        const long col = c < np ? c : np - 1;
        const long R = (col + 1) * 2 * d - 1;
        xo[v] = X[(DOWN ? R - d : R) * na + icl];
        xs[v] = DOWN ? S[col * na + icl] : DELTA ? X[(R - d) * na + icl] : 0.0f;
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const float sum = acc[v];
        const long col = c0 + wr * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
        if (col >= np || i >= na) continue;
        const long R = (col + 1) * 2 * d - 1;
        if (!DOWN) {
            X[R * na + i] = DELTA ? xo[v] + (xs[v] + sum) : xo[v] + sum;
        } else {
            const float y = DELTA ? xo[v] + (xs[v] + sum) : sum + xo[v];
            X[(R - d) * na + i] = xs[v];
            X[R * na + i] = y;
            if (Sn) {
                if (2 * col < npn) Sn[(2 * col) * na + i] = xs[v];
                if (2 * col + 1 < npn) Sn[(2 * col + 1) * na + i] = y;
            }
        }
    }
}

template <typename TT, typename TX, bool DOWN, bool DELTA, bool SMALL>
__global__ __launch_bounds__(256) void k_level_ks(int na, int ldp, long d, long np,
                                                  const TT* __restrict__ PT,
                                                  const TX* __restrict__ S, TX* __restrict__ X,
                                                  TX* __restrict__ Sn, long npn)
{
    d_level_ks<TT, TX, DOWN, DELTA, SMALL>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, ldp, d, np, PT, S, X, Sn, npn);
}

template <typename TX>
__device__ __forceinline__ void d_gather_rt(const VBlock vb, int na, long d, long np, const TX* __restrict__ X, TX* __restrict__ S)
{
    const long total = np * na;
    for (long e = (long)vb.x * blockDim.x + threadIdx.x; e < total;
         e += (long)vb.gx * blockDim.x) {
        const long c = e / na;
        const int j = (int)(e - c * na);
        S[e] = X[((c + 1) * 2 * d - 1) * na + j];
    }
}

template <typename TX>
__global__ void k_gather_rt(int na, long d, long np, const TX* __restrict__ X, TX* __restrict__ S)
{
    d_gather_rt<TX>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, d, np, X, S);
}

template <typename TT, typename TX>
void level_ks(hipStream_t s, bool down, int na, int ldp, long d, long np, const TT* PT, TX* X,
              TX* S, bool delta, bool gathered, TX* Sn, long npn)
{
    if (np <= 0) return;
    if (sizeof(TT) == 4 && ldp % 32 != 0)
        throw std::invalid_argument("level_ks: FP32 table must be the padded MFMA table");
    // FP32: 16 x 16 tiles while 32 x 32 ones would make fewer than 1024 work-groups
    const bool small = sizeof(TT) == 4 && ((np + 31) / 32) * ((na + 31) / 32) < 1024;
    const int T = sizeof(TT) == 4 && !small ? 32 : 16;
    dim3 g((unsigned)((np + T - 1) / T), (unsigned)((na + T - 1) / T)), b(256);
    if (down) {
        if (!S) throw std::invalid_argument("level_ks: the down-sweep needs the copy buffer");
        if (Sn && (Sn == S || npn < 0 || npn > 2 * np))
            throw std::invalid_argument("level_ks: bad next-level copy");
        if (!gathered)
            hipLaunchKernelGGL((k_gather_rt<TX>), dim3(blocks_for(np * na, 256, 4096)), dim3(256),
                               0, s, na, d, np, X, S);
    }
    auto go = [&](auto dn, auto dl, auto sm) {
        constexpr bool DN = decltype(dn)::value, DL = decltype(dl)::value;
        constexpr bool SM = decltype(sm)::value && sizeof(TT) == 4;
        hipLaunchKernelGGL((k_level_ks<TT, TX, DN, DL, SM>), g, b, 0, s, na, ldp, d, np, PT, S, X,
                           DN ? Sn : nullptr, DN ? npn : 0L);
    };
    using T0 = std::false_type;
    using T1 = std::true_type;
    if (down) {
        if (delta)
            small ? go(T1{}, T1{}, T1{}) : go(T1{}, T1{}, T0{});
        else
            small ? go(T1{}, T0{}, T1{}) : go(T1{}, T0{}, T0{});
    } else {
        if (delta)
            small ? go(T0{}, T1{}, T1{}) : go(T0{}, T1{}, T0{});
        else
            small ? go(T0{}, T0{}, T1{}) : go(T0{}, T0{}, T0{});
    }
    check_launch("level_ks");
}

template void level_ks<float, float>(hipStream_t, bool, int, int, long, long, const float*, float*,
                                     float*, bool, bool, float*, long);
template void level_ks<double, float>(hipStream_t, bool, int, int, long, long, const double*,
                                      float*, float*, bool, bool, float*, long);
template void level_ks<double, double>(hipStream_t, bool, int, int, long, long, const double*,
                                       double*, double*, bool, bool, double*, long);

// ------------------------------------------------------- top of the tree
// Levels [l0, L) of the up-sweep, the root and the same levels of the
// down-sweep in ONE work-group of 1024 threads: the top of the tree holds few
// pairs per level, so one work-group does the work of 2(L - l0) + 1 launches
// with only barriers between levels.  Levels >= l64 use the FP64 table.
// Thread t owns state i = t mod rowThreads and the column group t / rowThreads.
constexpr int kTopThreads = 1024;
constexpr int kTopMaxCols = 4;

// One level's update of cc staged columns for the rows of thread (i0, g).
template <typename TX, typename TA>
__device__ __forceinline__ void top_update(bool down, int na, long d, long c0, int cc, int i0,
                                           int rowThreads, int g, int groups,
                                           const TA* __restrict__ P, const double* Bsh, TX* X)
{
    for (int i = i0; i < na; i += rowThreads) {
        TA acc[kTopMaxCols];
#pragma unroll
        for (int q = 0; q < kTopMaxCols; ++q) acc[q] = 0;
        int j = 0;
        for (; j + 4 <= na; j += 4) {
            TA pj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pj[u] = P[(long)(j + u) * na + i];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < kTopMaxCols; ++q) {
                    const int c = g + q * groups;
                    if (c < cc) acc[q] = fma(pj[u], (TA)Bsh[c * na + j + u], acc[q]);
                }
        }
        for (; j < na; ++j) {
            const TA pj = P[(long)j * na + i];
#pragma unroll
            for (int q = 0; q < kTopMaxCols; ++q) {
                const int c = g + q * groups;
                if (c < cc) acc[q] = fma(pj, (TA)Bsh[c * na + j], acc[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < kTopMaxCols; ++q) {
            const int c = g + q * groups;
            if (c >= cc) continue;
            const long R = (c0 + c + 1) * 2 * d - 1, Lp = R - d;
            if (!down) {
                X[R * na + i] = (TX)((TA)X[R * na + i] + acc[q]);
            } else {
                const TA xl = (TA)X[Lp * na + i];
                X[Lp * na + i] = (TX)Bsh[c * na + i];
                X[R * na + i] = (TX)(acc[q] + xl);
            }
        }
    }
}

// Update for na >= 16 on the FP64 matrix cores: Out[i][c] = sum_j P[i][j] S[c][j]
// for the cc staged columns, A = P from the transposed FP64 table (lane
// l: i = 16 ib + l % 16, j = 4 k + l / 16) and B = S from LDS (pitch SP, odd).
// The (16-row block, 16-column tile) items are dealt to the waves in turn, so
// that a narrow state (one to three row blocks) still gives every wave a tile;
// each wave issues the loads of 8 k-steps before their MFMAs.
constexpr int kTopMfmaCols = 16;

// columns staged per pass of the matrix-core top: one 16-column tile per wave
// over the state's 16-row blocks, within 48 KiB of LDS
__host__ __device__ constexpr int top_mfma_cols(int na)
{
    const int nb = (na + 15) / 16, SP = na + 1;
    const int tiles = nb >= kTopThreads / 64 ? 1 : (kTopThreads / 64) / nb;
    const int cc = kTopMfmaCols * tiles;
    return cc < 6144 / SP ? cc : 6144 / SP;
}

template <typename TX>
__device__ __forceinline__ void top_update_mfma(bool down, int na, int SP, long d, long c0, int cc,
                                                const double* __restrict__ PT, const double* Ssh,
                                                TX* X)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
    const int nb = (na + 15) / 16, ks = (na + 3) / 4, ct = (cc + 15) / 16;
    for (int item = w; item < nb * ct; item += kTopThreads / 64) {
        const int ib = item % nb, cb = (item / nb) * 16;
        const int ia = ib * 16 + l16;
        doublex4 acc = {0.0, 0.0, 0.0, 0.0};
        constexpr int U = 8;
        for (int k = 0; k < ks; k += U) {
            double a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = (k + u) * 4 + q4;
                a[u] = (j < na && ia < na) ? PT[(long)j * na + ia] : 0.0;
                b[u] = (j < na && cb + l16 < cc) ? Ssh[(cb + l16) * SP + j] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
        }
        // D[i][c] is in lane c + 16 (i % 4), register i / 4
        const int c = cb + l16;
        if (c >= cc) continue;
        const long R = (c0 + c + 1) * 2 * d - 1, Lp = R - d;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = ib * 16 + 4 * v + q4;
            if (i >= na) continue;
            if (!down) {
                X[R * na + i] = (TX)((double)X[R * na + i] + acc[v]);
            } else {
                const double xl = (double)X[Lp * na + i];
                X[Lp * na + i] = (TX)Ssh[c * SP + i];
                X[R * na + i] = (TX)(acc[v] + xl);
            }
        }
    }
}

template <typename TX, bool MFMA>
__device__ __forceinline__ void d_top(const VBlock vb, int na, int l0, int L, int l64, long nReal,
                                                     const float* __restrict__ PT32,
                                                     const double* __restrict__ PT64,
                                                     TX* __restrict__ X,
                                                     const double* __restrict__ seed,
                                                     double* __restrict__ tot, int CC)
{
    extern __shared__ double Bsh[];   // CC x SP staged source columns
    const int t = threadIdx.x;
    const int SP = MFMA ? na + 1 : na;
    const int rowThreads = na < kTopThreads ? na : kTopThreads;
    const int groups = kTopThreads / rowThreads;
    const int i0 = t % rowThreads, g = t / rowThreads;
    const long na2 = (long)na * na;
    for (int pass = 0; pass < 2; ++pass) {
        const bool down = pass == 1;
        for (int li = 0; li < L - l0; ++li) {
            const int l = down ? L - 1 - li : l0 + li;
            const long d = 1L << l;
            const long np = (nReal + (2L << l) - 1) / (2L << l);
            for (long c0 = 0; c0 < np; c0 += CC) {
                const int cc = (np - c0) < CC ? (int)(np - c0) : CC;
                for (int e = t; e < cc * na; e += kTopThreads) {
                    const int c = e / na, j = e % na;
                    const long R = (c0 + c + 1) * 2 * d - 1;
                    Bsh[c * SP + j] = (double)X[(down ? R : R - d) * na + j];
                }
                __syncthreads();
                if (MFMA) {
                    top_update_mfma<TX>(down, na, SP, d, c0, cc, PT64 + (long)l * na2, Bsh, X);
                } else if (g < groups) {
                    if (l >= l64)
                        top_update<TX, double>(down, na, d, c0, cc, i0, rowThreads, g, groups,
                                               PT64 + (long)l * na2, Bsh, X);
                    else
                        top_update<TX, float>(down, na, d, c0, cc, i0, rowThreads, g, groups,
                                              PT32 + (long)l * na2, Bsh, X);
                }
                __syncthreads();
            }
        }
        if (!down) {
            const long last = (1L << L) - 1;
            for (int i = t; i < na; i += kTopThreads) {
                if (tot) tot[i] = (double)X[last * na + i];
                X[last * na + i] = (TX)seed[i];
            }
            __syncthreads();
        }
    }
}

template <typename TX, bool MFMA>
__global__ __launch_bounds__(kTopThreads) void k_top(int na, int l0, int L, int l64, long nReal,
                                                     const float* __restrict__ PT32,
                                                     const double* __restrict__ PT64,
                                                     TX* __restrict__ X,
                                                     const double* __restrict__ seed,
                                                     double* __restrict__ tot, int CC)
{
    d_top<TX, MFMA>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, l0, L, l64, nReal, PT32, PT64, X, seed, tot, CC);
}

template <typename TX>
void top(hipStream_t s, int na, int l0, int L, int l64, long nReal, const float* PT32,
         const double* PT64, TX* X, const double* seed, double* tot)
{
    // columns per pass: a 16-column MFMA tile per wave for na >= 16, else what
    // the threads can hold (groups x kTopMaxCols); at most 48 KiB of LDS staged
    const int rowThreads = na < kTopThreads ? na : kTopThreads;
    const int groups = kTopThreads / rowThreads;
    const int SP = na >= 16 ? na + 1 : na;
    int CC = na >= 16 ? top_mfma_cols(na) : std::min(groups * kTopMaxCols, 6144 / na);
    if (CC < 1) CC = 1;
    size_t lds = (size_t)CC * SP * sizeof(double);
    if (na >= 16)
        hipLaunchKernelGGL((k_top<TX, true>), dim3(1), dim3(kTopThreads), lds, s, na, l0, L, l64,
                           nReal, PT32, PT64, X, seed, tot, CC);
    else
        hipLaunchKernelGGL((k_top<TX, false>), dim3(1), dim3(kTopThreads), lds, s, na, l0, L, l64,
                           nReal, PT32, PT64, X, seed, tot, CC);
    check_launch("top");
}

template void top<float>(hipStream_t, int, int, int, int, long, const float*, const double*,
                         float*, const double*, double*);
template void top<double>(hipStream_t, int, int, int, int, long, const float*, const double*,
                          double*, const double*, double*);

// -------------------------------------------------------------------- root
template <typename TX, typename TS>
__device__ __forceinline__ void d_root(const VBlock vb, int na, long last, const TS* __restrict__ seed, TX* __restrict__ X,
                       double* __restrict__ tot)
{
    for (int i = threadIdx.x; i < na; i += blockDim.x) {
        if (tot) tot[i] = (double)X[last * na + i];
        X[last * na + i] = (TX)seed[i];
    }
}

template <typename TX, typename TS>
__global__ void k_root(int na, long last, const TS* __restrict__ seed, TX* __restrict__ X,
                       double* __restrict__ tot)
{
    d_root<TX, TS>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, na, last, seed, X, tot);
}

template <typename TX, typename TS>
void root(hipStream_t s, int na, long last, const TS* seed, TX* X, double* tot)
{
    hipLaunchKernelGGL((k_root<TX, TS>), dim3(1), dim3(256), 0, s, na, last, seed, X, tot);
    check_launch("root");
}

template void root<float, double>(hipStream_t, int, long, const double*, float*, double*);
template void root<double, double>(hipStream_t, int, long, const double*, double*, double*);

__global__ void k_carry(int na, const double* __restrict__ PT, const double* __restrict__ seed,
                        const double* __restrict__ tot, double* __restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < na; ++j) acc = fma(PT[(long)j * na + i], seed[j], acc);
        out[i] = acc + tot[i];
    }
}

void carry(hipStream_t s, int na, const double* P, const double* seed, const double* tot,
           double* seedOut)
{
    hipLaunchKernelGGL(k_carry, dim3((na + 255) / 256), dim3(256), 0, s, na, P, seed, tot, seedOut);
    check_launch("carry");
}

// ----------------------------------------------------------------- outputs
// A work-group takes epb elements at a time: their state blocks (n values at
// offset oX*n of each element) and input rows are loaded coalesced into LDS,
// with C^T and D^T staged once; then 2^tpeLog threads per element form its
// outputs from LDS.  The sum runs over c in order, without contraction (the
// reference's SystemOutput arithmetic in the plain path).
template <typename TX, typename TU>
__global__ __launch_bounds__(256) void k_outputs(int n, int m, int p, int na, int oX, bool zC,
                                                 bool zD, long N, int tpeLog, int epb,
                                                 const float* __restrict__ CT,
                                                 const float* __restrict__ DT,
                                                 const TX* __restrict__ X,
                                                 const TU* __restrict__ U, long sO, long oO,
                                                 float* __restrict__ Y)
{
#pragma clang fp contract(off)
    using TA = decltype(TX(0) * TU(0));
    extern __shared__ unsigned char out_smem[];
    const int nOut = (zC && zD) ? n : p;
    const int useC = zC ? 0 : 1, useD = (zD || m <= 0) ? 0 : 1;
    const int xp = n + 1;                                   // odd pitch
    // [epb][xp] states, [epb][m] inputs, [n][p] C^T, [m][p] D^T; 16-byte aligned parts
    size_t off = 0;
    TX* xs = reinterpret_cast<TX*>(out_smem);
    off += ((size_t)epb * xp * sizeof(TX) + 15) & ~(size_t)15;
    TU* us = reinterpret_cast<TU*>(out_smem + off);
    off += ((size_t)epb * m * useD * sizeof(TU) + 15) & ~(size_t)15;
    float* cs = reinterpret_cast<float*>(out_smem + off);
    float* ds = cs + (size_t)n * p * useC;
    for (int t = threadIdx.x; t < n * p * useC; t += 256) cs[t] = CT[t];
    for (int t = threadIdx.x; t < m * p * useD; t += 256) ds[t] = DT[t];
    const int tpe = 1 << tpeLog;
    const int le = threadIdx.x >> tpeLog, lane = threadIdx.x & (tpe - 1);
    for (long e0 = (long)blockIdx.x * epb; e0 < N; e0 += (long)gridDim.x * epb) {
        const int cnt = (N - e0) < epb ? (int)(N - e0) : epb;
        __syncthreads();
        for (int t = threadIdx.x; t < cnt * n; t += 256) {
            const int q = t / n, c = t - q * n;
            xs[q * xp + c] = X[(e0 + q) * na + (long)oX * n + c];
        }
        for (int t = threadIdx.x; t < cnt * m * useD; t += 256) {
            const int q = t / m, c = t - q * m;
            us[t] = U[((e0 + q) * sO + oO) * m + c];
        }
        __syncthreads();
        if (le >= cnt) continue;
        const TX* x = xs + le * xp;
        float* y = Y + (e0 + le) * nOut;
        for (int j = lane; j < nOut; j += tpe) {
            if (zC && zD) {
                y[j] = (float)x[j];
                continue;
            }
            TA v = 0;
            if (useC)
                for (int c = 0; c < n; ++c) v = madd((TA)cs[c * p + j], (TA)x[c], v);
            if (useD) {
                const TU* u = us + le * m;
                for (int c = 0; c < m; ++c) v = madd((TA)ds[c * p + j], (TA)u[c], v);
            }
            y[j] = (float)v;
        }
    }
}

// The same outputs with C^T and D^T read from global memory (and L2), for
// output matrices too large to stage in LDS.
template <typename TX, typename TU>
__global__ __launch_bounds__(256) void k_outputs_direct(int n, int m, int p, int na, int oX,
                                                        bool zC, bool zD, long N, int tpeLog,
                                                        const float* __restrict__ CT,
                                                        const float* __restrict__ DT,
                                                        const TX* __restrict__ X,
                                                        const TU* __restrict__ U, long sO,
                                                        long oO, float* __restrict__ Y)
{
#pragma clang fp contract(off)
    using TA = decltype(TX(0) * TU(0));
    const int nOut = (zC && zD) ? n : p;
    const int tpe = 1 << tpeLog;
    const int epb = 256 >> tpeLog;
    const int lane = threadIdx.x & (tpe - 1);
    for (long e = (long)blockIdx.x * epb + (threadIdx.x >> tpeLog); e < N;
         e += (long)gridDim.x * epb) {
        const TX* x = X + e * na + (long)oX * n;
        float* y = Y + e * nOut;
        for (int j = lane; j < nOut; j += tpe) {
            if (zC && zD) {
                y[j] = (float)x[j];
                continue;
            }
            TA v = 0;
            if (!zC)
                for (int c = 0; c < n; ++c) {
                    TA pr = (TA)CT[(long)c * p + j] * (TA)x[c];
                    v = v + pr;
                }
            if (!zD && m > 0) {
                const TU* u = U + (e * sO + oO) * m;
                for (int c = 0; c < m; ++c) {
                    TA pr = (TA)DT[(long)c * p + j] * (TA)u[c];
                    v = v + pr;
                }
            }
            y[j] = (float)v;
        }
    }
}

// Outputs with one thread per element (a work-group of kOutLanes elements), in
// FP64: the elements' state blocks are staged in LDS (coalesced rows, odd
// pitch); C^T and D^T, FP64 with rows padded with zeros to pp (a multiple of
// 8) columns, are read at addresses that are the same for every lane (scalar
// loads of eight values), and each thread accumulates eight of its element's
// outputs at a time in registers.  Sums over c in order, C terms then D terms,
// as k_outputs.
constexpr int kOutLanes = 64;

template <typename TX, typename TU>
__device__ __forceinline__ void d_outputs_lanes(const VBlock vb, int n, int m, int p, int pp, int na, int oX, bool useC, bool useD, long N,
    const double* __restrict__ CT, const double* __restrict__ DT, const TX* __restrict__ X,
    const TU* __restrict__ U, long sO, long oO, float* __restrict__ Y)
{
    extern __shared__ unsigned char outl_smem[];
    TX* xs = reinterpret_cast<TX*>(outl_smem);   // [kOutLanes][n + 1]
    const int xp = n + 1;
    const long stride = (long)vb.gx * kOutLanes;
    // With n <= kOutLanes, thread c holds state c of every element of a block:
    // the next block's states are loaded into registers while this one computes.
    // (FP32 states only: FP64 ones would take twice the registers)
    constexpr int XV = sizeof(TX) == 4 ? kOutLanes : 16;
    const bool ahead = useC && n <= kOutLanes && sizeof(TX) == 4;
    TX xv[XV];
    auto load_x = [&](long e0) {
        const int cnt = (N - e0) < kOutLanes ? (int)(N - e0) : kOutLanes;
        const int c = threadIdx.x < n ? (int)threadIdx.x : n - 1;
#pragma unroll
        for (int u = 0; u < XV; ++u) {
            const int q = u < cnt ? u : cnt - 1;
            xv[u] = X[(e0 + q) * na + (long)oX * n + c];
        }
    };
    if (ahead && (long)vb.x * kOutLanes < N) load_x((long)vb.x * kOutLanes);
    for (long e0 = (long)vb.x * kOutLanes; e0 < N; e0 += stride) {
        const int cnt = (N - e0) < kOutLanes ? (int)(N - e0) : kOutLanes;
        __syncthreads();
        if (ahead) {
            if ((int)threadIdx.x < n)
#pragma unroll
                for (int u = 0; u < XV; ++u)
                    if (u < cnt) xs[u * xp + threadIdx.x] = xv[u];
        } else if (useC) {
            // sixteen elements' rows in flight per thread
            for (int c = threadIdx.x; c < n; c += kOutLanes)
                for (int q0 = 0; q0 < cnt; q0 += 16) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        const int q = q0 + u < cnt ? q0 + u : cnt - 1;
                        xv[u] = X[(e0 + q) * na + (long)oX * n + c];
                    }
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        if (q0 + u < cnt) xs[(q0 + u) * xp + c] = xv[u];
                }
        }
        __syncthreads();
        if (ahead && e0 + stride < N) load_x(e0 + stride);
        if ((int)threadIdx.x >= cnt) continue;
        const long e = e0 + threadIdx.x;
        const TX* x = xs + threadIdx.x * xp;
        const TU* u = U + (e * sO + oO) * m;
        float* y = Y + e * p;
        for (int jb = 0; jb < p; jb += 8) {
            double v[8];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) v[jj] = 0.0;
            if (useC) {
                // two rows of C^T per batch, so their scalar loads are in
                // flight together
                int c = 0;
                for (; c + 2 <= n; c += 2) {
                    double cr[2][8], xc[2];
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        xc[q] = (double)x[c + q];
#pragma unroll
                        for (int jj = 0; jj < 8; ++jj) cr[q][jj] = CT[(long)(c + q) * pp + jb + jj];
                    }
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int jj = 0; jj < 8; ++jj) v[jj] = fma(cr[q][jj], xc[q], v[jj]);
                }
                for (; c < n; ++c) {
                    const double xc = (double)x[c];
                    const double* cr = CT + (long)c * pp + jb;
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) v[jj] = fma(cr[jj], xc, v[jj]);
                }
            }
            if (useD)
                for (int c0 = 0; c0 < m; c0 += 8) {
                    double uv[8];   // eight inputs in flight (indices past m repeat m - 1)
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) uv[cc] = (double)u[c0 + cc < m ? c0 + cc : m - 1];
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        if (c0 + cc >= m) break;
                        const double* dr = DT + (long)(c0 + cc) * pp + jb;
#pragma unroll
                        for (int jj = 0; jj < 8; ++jj) v[jj] = fma(dr[jj], uv[cc], v[jj]);
                    }
                }
#pragma unroll
            for (int jj = 0; jj < 8; ++jj)
                if (jb + jj < p) y[jb + jj] = (float)v[jj];
        }
    }
}

template <typename TX, typename TU>
__global__ __launch_bounds__(kOutLanes) void k_outputs_lanes(int n, int m, int p, int pp, int na, int oX, bool useC, bool useD, long N,
    const double* __restrict__ CT, const double* __restrict__ DT, const TX* __restrict__ X,
    const TU* __restrict__ U, long sO, long oO, float* __restrict__ Y)
{
    d_outputs_lanes<TX, TU>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, m, p, pp, na, oX, useC, useD, N, CT, DT, X, U, sO, oO, Y);
}

// The same outputs for wide states: the states go through LDS in chunks of
// kOutKc per element (64 x (kOutKc + 1) words per work-group, whatever n),
// so many one-wave work-groups share a CU, and each lane keeps JG FP64 sums
// (a group of outputs) over the chunks.  C^T and D^T rows are read with
// scalar loads as above; the states are read coalesced (32 consecutive
// states of two elements per load).
constexpr int kOutKc = 32;

template <typename TX, typename TU, int JG>
__device__ __forceinline__ void d_outputs_kc(const VBlock vb, int n, int m, int p, int pp, int na, int oX, bool useC, bool useD, long N,
    const double* __restrict__ CT, const double* __restrict__ DT, const TX* __restrict__ X,
    const TU* __restrict__ U, long sO, long oO, float* __restrict__ Y)
{
    __shared__ TX xs[kOutLanes * (kOutKc + 1)];
    constexpr int XP = kOutKc + 1;
    const int t = threadIdx.x, cl = t & (kOutKc - 1), r0 = t / kOutKc;
    constexpr int RPL = kOutLanes / 2;    // staging loads per lane (two elements each)
    static_assert(kOutLanes == 2 * kOutKc, "two elements per staging load");
    const long stride = (long)vb.gx * kOutLanes;
    for (long e0 = (long)vb.x * kOutLanes; e0 < N; e0 += stride) {
        const int cnt = (N - e0) < kOutLanes ? (int)(N - e0) : kOutLanes;
        const long e = e0 + (t < cnt ? t : cnt - 1);
        const TU* u = U + (e * sO + oO) * m;
        for (int jb = 0; jb < pp; jb += JG) {
            double v[JG];
#pragma unroll
            for (int jj = 0; jj < JG; ++jj) v[jj] = 0.0;
            if (useC)
                for (int kc = 0; kc < n; kc += kOutKc) {
                    const int kn = (n - kc) < kOutKc ? n - kc : kOutKc;
                    TX xv[RPL];
#pragma unroll
                    for (int q = 0; q < RPL; ++q) {
                        const int r = r0 + 2 * q;
                        const long er = e0 + (r < cnt ? r : cnt - 1);
                        xv[q] = cl < kn ? X[er * na + (long)oX * n + kc + cl] : (TX)0;
                    }
                    __syncthreads();              // the previous chunk has been read
#pragma unroll
                    for (int q = 0; q < RPL; ++q) xs[(r0 + 2 * q) * XP + cl] = xv[q];
                    __syncthreads();
                    const TX* x = xs + t * XP;
                    // RB rows of C^T per batch: at most 32 doubles in SGPRs
                    constexpr int RB = JG > 16 ? 1 : 32 / JG;
                    int c = 0;
#pragma unroll 1
                    for (; c + RB <= kn; c += RB) {
                        double cr[RB][JG], xc[RB];
#pragma unroll
                        for (int q = 0; q < RB; ++q) {
                            xc[q] = (double)x[c + q];
#pragma unroll
                            for (int jj = 0; jj < JG; ++jj)
                                cr[q][jj] = CT[(long)(kc + c + q) * pp + jb + jj];
                        }
#pragma unroll
                        for (int q = 0; q < RB; ++q)
#pragma unroll
                            for (int jj = 0; jj < JG; ++jj) v[jj] = fma(cr[q][jj], xc[q], v[jj]);
                    }
#pragma unroll 1
                    for (; c < kn; ++c) {
                        const double xc = (double)x[c];
                        const double* cr = CT + (long)(kc + c) * pp + jb;
#pragma unroll
                        for (int jj = 0; jj < JG; ++jj) v[jj] = fma(cr[jj], xc, v[jj]);
                    }
                }
            if (useD)
                for (int c0 = 0; c0 < m; c0 += 8) {
                    double uv[8];   // eight inputs in flight (indices past m repeat m - 1)
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) uv[cc] = (double)u[c0 + cc < m ? c0 + cc : m - 1];
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        if (c0 + cc >= m) break;
                        const double* dr = DT + (long)(c0 + cc) * pp + jb;
#pragma unroll
                        for (int jj = 0; jj < JG; ++jj) v[jj] = fma(dr[jj], uv[cc], v[jj]);
                    }
                }
            if (t < cnt) {
                float* y = Y + e * p;
#pragma unroll
                for (int jj = 0; jj < JG; ++jj)
                    if (jb + jj < p) y[jb + jj] = (float)v[jj];
            }
        }
    }
}

template <typename TX, typename TU, int JG>
__global__ __launch_bounds__(kOutLanes) void k_outputs_kc(int n, int m, int p, int pp, int na, int oX, bool useC, bool useD, long N,
    const double* __restrict__ CT, const double* __restrict__ DT, const TX* __restrict__ X,
    const TU* __restrict__ U, long sO, long oO, float* __restrict__ Y)
{
    d_outputs_kc<TX, TU, JG>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, m, p, pp, na, oX, useC, useD, N, CT, DT, X, U, sO, oO, Y);
}

// The same outputs on the FP64 matrix cores (v_mfma_f64_16x16x4f64), as the
// product of the elements' state blocks (FP32, converted exactly) and inputs
// with C^T and D^T: a work-group of four waves takes 64 elements, whose state
// blocks go through LDS in chunks of kOutKcM states (coalesced rows); each wave
// forms 16 of the elements times up to kOutMfmaTiles 16-output tiles, K over
// the n states, then the m inputs (read directly, FP64).  Fragments:
//   A (16 elements x 4 k): lane l holds element e0 + l%16, k = k0 + l/16;
//   B (4 k x 16 outputs):  lane l holds CT[k0 + l/16][j0 + l%16];
//   D: lane l, register v holds element e0 + 4v + l/16, output j0 + l%16.
// Output columns past pp (the padding of C^T and D^T) and k past n or m are
// read clamped (their A values are zero); elements past N repeat the last one
// and are not stored.
constexpr int kOutMfmaTiles = 4, kOutKcM = 32;

__device__ __forceinline__ void d_outputs_mfma(const VBlock vb, int n, int m, int p, int pp,
                                               int na, int oX, bool useC, bool useD, long N,
                                               const double* __restrict__ CT,
                                               const double* __restrict__ DT,
                                               const float* __restrict__ X,
                                               const double* __restrict__ U, long sO, long oO,
                                               float* __restrict__ Y)
{
    __shared__ float xs[64][kOutKcM + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, k4 = lane >> 4;
    const int jtiles = (p + 15) / 16;
    const int sel = tid >> 2, sb = (tid & 3) * 8;   // staging: element, first of 8 states
    for (long eb = vb.x * 64; eb < N; eb += vb.gx * 64) {
        const int cnt = N - eb < 64 ? (int)(N - eb) : 64;
        const long e0 = eb + 16 * wave;
        const long ea = eb + (16 * wave + l16 < cnt ? 16 * wave + l16 : cnt - 1);
        const float* xstage = X + (eb + (sel < cnt ? sel : cnt - 1)) * na + (long)oX * n;
        const double* ur = U + (ea * sO + oO) * m;
        for (int jt0 = 0; jt0 < jtiles; jt0 += kOutMfmaTiles) {
            const int nt = jtiles - jt0 < kOutMfmaTiles ? jtiles - jt0 : kOutMfmaTiles;
            doublex4 acc[kOutMfmaTiles];
            int jb[kOutMfmaTiles];
#pragma unroll
            for (int t = 0; t < kOutMfmaTiles; ++t) {
                acc[t] = doublex4{0.0, 0.0, 0.0, 0.0};
                const int j = (jt0 + t) * 16 + l16;
                jb[t] = j < pp ? j : pp - 1;
            }
            if (useC) {
                for (int kc0 = 0; kc0 < n; kc0 += kOutKcM) {
                    float v[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int k = kc0 + sb + q;
                        v[q] = k < n ? xstage[k] : 0.0f;
                    }
                    __syncthreads();                  // the previous chunk has been read
#pragma unroll
                    for (int q = 0; q < 8; ++q) xs[sel][sb + q] = v[q];
                    __syncthreads();
                    const int kn = n - kc0 < kOutKcM ? n - kc0 : kOutKcM;
                    for (int ks = 0; ks < kn; ks += 4) {
                        const double a = (double)xs[16 * wave + l16][ks + k4];
                        const int kc = kc0 + ks + k4 < n ? kc0 + ks + k4 : n - 1;
                        double b[kOutMfmaTiles];
#pragma unroll
                        for (int t = 0; t < kOutMfmaTiles; ++t)
                            b[t] = t < nt ? CT[(long)kc * pp + jb[t]] : 0.0;
#pragma unroll
                        for (int t = 0; t < kOutMfmaTiles; ++t)
                            if (t < nt)
                                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[t], acc[t], 0,
                                                                             0, 0);
                    }
                }
            }
            if (useD) {
                for (int k0 = 0; k0 < m; k0 += 4) {
                    const int k = k0 + k4, kc = k < m ? k : m - 1;
                    const double a = k < m ? ur[kc] : 0.0;
                    double b[kOutMfmaTiles];
#pragma unroll
                    for (int t = 0; t < kOutMfmaTiles; ++t)
                        b[t] = t < nt ? DT[(long)kc * pp + jb[t]] : 0.0;
#pragma unroll
                    for (int t = 0; t < kOutMfmaTiles; ++t)
                        if (t < nt)
                            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[t], acc[t], 0, 0,
                                                                         0);
                }
            }
#pragma unroll
            for (int t = 0; t < kOutMfmaTiles; ++t) {
                const int j = (jt0 + t) * 16 + l16;
                if (t >= nt || j >= p) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const long e = e0 + 4 * v + k4;
                    if (e < N) Y[e * p + j] = (float)acc[t][v];
                }
            }
        }
    }
}

__global__ void k_pad_cast(int rows, int cols, int pp, const float* __restrict__ in,
                           double* __restrict__ out)
{
    const int total = rows * pp;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int r = t / pp, c = t - r * pp;
        out[t] = c < cols ? (double)in[(long)r * cols + c] : 0.0;
    }
}

void pad_cast(hipStream_t s, int rows, int cols, int pp, const float* in, double* out)
{
    if (rows <= 0 || pp <= 0) return;
    hipLaunchKernelGGL(k_pad_cast, dim3(blocks_for((long)rows * pp, 256, 1024)), dim3(256), 0, s,
                       rows, cols, pp, in, out);
    check_launch("pad_cast");
}

// Outputs as tiles of 64 elements x 16 JT outputs of the product
// Y[e][j] = sum_k [x_e | u_e][k] [C^T ; D^T][k][j] on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64): four waves, each 16 elements x JT output tiles (one
// A fragment feeds JT products).  K runs over the n states (with C) in slices
// of KB staged in LDS in the states' own type (FP32 states convert exactly at
// the product; a slice of 32 FP32 states costs the LDS of 16 FP64 ones, twice
// the bytes in flight per load), then over the m inputs (with D) in FP64
// slices of 16; both double buffered.  Elements past N read zero and are not
// stored; output columns past pp (the padding of C^T and D^T) read zero.
// Grid: x = output tile, y = element tile.
constexpr int kOG = 64, kOGU = 16;

template <typename TX, typename TU, int JT>
__device__ __forceinline__ void d_outputs_gemm(const VBlock vb, int n, int m, int p, int pp,
                                               int na, int oX, bool useC, bool useD, long N,
                                               const double* __restrict__ CT,
                                               const double* __restrict__ DT,
                                               const TX* __restrict__ X,
                                               const TU* __restrict__ U, long sO, long oO,
                                               float* __restrict__ Y)
{
    constexpr int OW = 16 * JT;                      // outputs per tile
    constexpr int KB = sizeof(TX) == 4 ? 32 : 16;    // states per slice
    constexpr int AQ = KB / 4;                       // A values per thread and slice
    constexpr int BQ = KB * OW / 256;                // C^T values per thread and slice
    constexpr int BQU = OW / 16;                     // D^T values per thread and slice
    // the states' slices, then (the same LDS) the inputs' slices
    constexpr int AXW = (2 * kOG * (KB + 1) * (int)sizeof(TX) + 7) / 8;   // in doubles
    constexpr int X_WORDS = AXW + 2 * KB * (OW + 1);
    constexpr int U_WORDS = 2 * kOG * (kOGU + 1) + 2 * kOGU * (OW + 1);
    __shared__ double smem[X_WORDS > U_WORDS ? X_WORDS : U_WORDS];
    TX(*Ax)[kOG][KB + 1] = reinterpret_cast<TX(*)[kOG][KB + 1]>(smem);
    double(*Bx)[KB][OW + 1] = reinterpret_cast<double(*)[KB][OW + 1]>(smem + AXW);
    double(*Au)[kOG][kOGU + 1] = reinterpret_cast<double(*)[kOG][kOGU + 1]>(smem);
    double(*Bu)[kOGU][OW + 1] =
        reinterpret_cast<double(*)[kOGU][OW + 1]>(smem + 2 * kOG * (kOGU + 1));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
    const long e0 = vb.y * kOG;
    const int j0 = (int)vb.x * OW;
    const int kC = useC ? n : 0, kU = useD ? m : 0;
    const int ar = tid >> 2;                         // A slice: element
    const long ea = e0 + ar;
    const bool realA = ea < N;
    const TX* xe = X + ea * na + (long)oX * n;
    const TU* ue = U + (ea * sO + oO) * m;
    doublex4 acc[JT];
#pragma unroll
    for (int t = 0; t < JT; ++t) acc[t] = doublex4{0.0, 0.0, 0.0, 0.0};
    // C x: states in slices of KB
    if (kC > 0) {
        const int ak = (tid & 3) * AQ;
        const int bk = tid / (OW / BQ), bj = (tid % (OW / BQ)) * BQ;
        TX ra[AQ];
        double rb[BQ];
        auto load = [&](int k0) {
            const int kf = k0 + ak;
            const TX* src = xe + kf;
            // FP32 states in range: AQ / 4 loads of 16 bytes where aligned, else
            // AQ / 2 loads of 8 (a quarter or half of the load instructions)
            const bool whole = sizeof(TX) == 4 && AQ % 4 == 0 && realA && kf + AQ <= kC;
            const uintptr_t al = reinterpret_cast<uintptr_t>(src) & 15;
            if (whole && al == 0) {
#pragma unroll
                for (int q = 0; q < AQ; q += 4) {
                    const float4 v = *reinterpret_cast<const float4*>(src + q);
                    ra[q] = (TX)v.x;
                    ra[q + 1] = (TX)v.y;
                    ra[q + 2] = (TX)v.z;
                    ra[q + 3] = (TX)v.w;
                }
            } else if (whole && (al & 7) == 0) {
#pragma unroll
                for (int q = 0; q < AQ; q += 2) {
                    const float2 v = *reinterpret_cast<const float2*>(src + q);
                    ra[q] = (TX)v.x;
                    ra[q + 1] = (TX)v.y;
                }
            } else {
#pragma unroll
                for (int q = 0; q < AQ; ++q) {
                    const int k = kf + q;
                    ra[q] = (realA && k < kC) ? xe[k] : (TX)0;
                }
            }
            const int kb = k0 + bk;
#pragma unroll
            for (int q = 0; q < BQ; ++q) {
                const int j = j0 + bj + q;
                rb[q] = (kb < kC && j < pp) ? CT[(long)kb * pp + j] : 0.0;
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int q = 0; q < AQ; ++q) Ax[buf][ar][ak + q] = ra[q];
#pragma unroll
            for (int q = 0; q < BQ; ++q) Bx[buf][bk][bj + q] = rb[q];
        };
        load(0);
        store(0);
        __syncthreads();
        int buf = 0;
        for (int k0 = 0; k0 < kC; k0 += KB) {
            const bool more = k0 + KB < kC;
            if (more) load(k0 + KB);
#pragma unroll
            for (int ks = 0; ks < KB; ks += 4) {
                const double a = (double)Ax[buf][wave * 16 + l16][ks + q4];
#pragma unroll
                for (int t = 0; t < JT; ++t)
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        a, Bx[buf][ks + q4][t * 16 + l16], acc[t], 0, 0, 0);
            }
            if (more) store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
    // D u: inputs in FP64 slices of 16
    if (kU > 0) {
        const int ak = (tid & 3) * 4;
        const int bk = tid >> 4, bj = (tid & 15) * BQU;
        double ra[4], rb[BQU];
        auto load = [&](int k0) {
            const TU* src = ue + k0 + ak;
            if (sizeof(TU) == 8 && realA && k0 + ak + 4 <= kU &&
                (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
                // four FP64 inputs in range: two loads of 16 bytes
#pragma unroll
                for (int q = 0; q < 4; q += 2) {
                    const double2 v = *reinterpret_cast<const double2*>(src + q);
                    ra[q] = (double)v.x;
                    ra[q + 1] = (double)v.y;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = k0 + ak + q;
                    ra[q] = (realA && k < kU) ? (double)ue[k] : 0.0;
                }
            }
            const int kb = k0 + bk;
#pragma unroll
            for (int q = 0; q < BQU; ++q) {
                const int j = j0 + bj + q;
                rb[q] = (kb < kU && j < pp) ? DT[(long)kb * pp + j] : 0.0;
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int q = 0; q < 4; ++q) Au[buf][ar][ak + q] = ra[q];
#pragma unroll
            for (int q = 0; q < BQU; ++q) Bu[buf][bk][bj + q] = rb[q];
        };
        load(0);
        store(0);
        __syncthreads();
        int buf = 0;
        for (int k0 = 0; k0 < kU; k0 += kOGU) {
            const bool more = k0 + kOGU < kU;
            if (more) load(k0 + kOGU);
#pragma unroll
            for (int ks = 0; ks < kOGU; ks += 4) {
                const double a = Au[buf][wave * 16 + l16][ks + q4];
#pragma unroll
                for (int t = 0; t < JT; ++t)
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        a, Bu[buf][ks + q4][t * 16 + l16], acc[t], 0, 0, 0);
            }
            if (more) store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
    // D[i][j] (element i, output j) is in lane j + 16 (i % 4), register i / 4
#pragma unroll
    for (int t = 0; t < JT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long e = e0 + wave * 16 + 4 * v + q4;
            const int j = j0 + t * 16 + l16;
            if (e < N && j < p) Y[e * p + j] = (float)acc[t][v];
        }
}

template <typename TX, typename TU, int JT>
__global__ __launch_bounds__(256) void k_outputs_gemm(int n, int m, int p, int pp, int na, int oX,
                                                      bool useC, bool useD, long N,
                                                      const double* __restrict__ CT,
                                                      const double* __restrict__ DT,
                                                      const TX* __restrict__ X,
                                                      const TU* __restrict__ U, long sO, long oO,
                                                      float* __restrict__ Y)
{
    d_outputs_gemm<TX, TU, JT>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, n, m,
                               p, pp, na, oX, useC, useD, N, CT, DT, X, U, sO, oO, Y);
}

// output tiles of 16, 32 or 64 outputs
static int outputs_gemm_width(int p) { return p <= 16 ? 1 : p <= 32 ? 2 : 4; }

template <typename TX, typename TU>
bool outputs_lanes(hipStream_t s, int n, int m, int p, int pp, int na, int oX, bool zC, bool zD,
                   long N, const double* CT, const double* DT, const TX* X, const TU* U, long sO,
                   long oO, float* Y)
{
    const bool useC = !zC, useD = !zD && m > 0;
    if (N <= 0 || p <= 0) return N <= 0;
    if (zC && zD) return false;   // the outputs are the states: k_outputs copies them
    if (pp % 8 != 0 || pp < p) throw std::invalid_argument("outputs_lanes: bad padding");
    if (useC && !(sizeof(TX) == 4 && n <= kOutLanes)) {
        // wide (or FP64) states: 64-element tiles on the FP64 matrix cores
        const int jt = outputs_gemm_width(p);
        const dim3 g2((unsigned)((p + 16 * jt - 1) / (16 * jt)), (unsigned)((N + kOG - 1) / kOG));
        auto go = [&](auto w) {
            constexpr int JT = decltype(w)::value;
            hipLaunchKernelGGL((k_outputs_gemm<TX, TU, JT>), g2, dim3(256), 0, s, n, m, p, pp, na,
                               oX, useC, useD, N, CT, DT, X, U, sO, oO, Y);
        };
        if (jt == 1)
            go(std::integral_constant<int, 1>());
        else if (jt == 2)
            go(std::integral_constant<int, 2>());
        else
            go(std::integral_constant<int, 4>());
        check_launch("outputs_gemm");
        return true;
    }
    // persistent work-groups (several blocks each)
    const long blocks = std::min<long>((N + kOutLanes - 1) / kOutLanes, 4096);
    const dim3 grid((unsigned)std::max<long>(blocks, 1));
    if (!(sizeof(TX) == 4 && n <= kOutLanes)) {
        // wide (or FP64) states: chunked through LDS, outputs in groups of JG
        // that divide the padded count
        auto go = [&](auto jg) {
            constexpr int JG = decltype(jg)::value;
            hipLaunchKernelGGL((k_outputs_kc<TX, TU, JG>), grid, dim3(kOutLanes), 0, s, n, m, p,
                               pp, na, oX, useC, useD, N, CT, DT, X, U, sO, oO, Y);
        };
        if (pp % 32 == 0)
            go(std::integral_constant<int, 32>());
        else if (pp % 24 == 0)
            go(std::integral_constant<int, 24>());
        else if (pp % 16 == 0)
            go(std::integral_constant<int, 16>());
        else
            go(std::integral_constant<int, 8>());
        check_launch("outputs_kc");
        return true;
    }
    const size_t lds = (size_t)kOutLanes * (n + 1) * sizeof(TX);
    hipLaunchKernelGGL((k_outputs_lanes<TX, TU>), grid, dim3(kOutLanes), lds, s, n, m, p, pp, na,
                       oX, useC, useD, N, CT, DT, X, U, sO, oO, Y);
    check_launch("outputs_lanes");
    return true;
}

template bool outputs_lanes<float, double>(hipStream_t, int, int, int, int, int, int, bool, bool,
                                           long, const double*, const double*, const float*,
                                           const double*, long, long, float*);
template bool outputs_lanes<double, double>(hipStream_t, int, int, int, int, int, int, bool,
                                            bool, long, const double*, const double*,
                                            const double*, const double*, long, long, float*);
template bool outputs_lanes<double, float>(hipStream_t, int, int, int, int, int, int, bool, bool,
                                           long, const double*, const double*, const double*,
                                           const float*, long, long, float*);

template <typename TX, typename TU>
void outputs(hipStream_t s, int n, int m, int p, int na, int oX, bool zC, bool zD, long N,
             const float* C, const float* D, const TX* X, const TU* U, long sO, long oO, float* Y)
{
    const int nOut = (zC && zD) ? n : p;
    if (N <= 0 || nOut <= 0) return;
    int tpeLog = 0;
    while ((1 << tpeLog) < nOut && tpeLog < 8) ++tpeLog;
    const bool useC = !zC, useD = !zD && m > 0;
    // elements per pass: one per thread group, as many as 48 KiB of LDS holds
    const size_t fixed = (useC ? (size_t)n * p * 4 : 0) + (useD ? (size_t)m * p * 4 : 0);
    const size_t per = (size_t)(n + 1) * sizeof(TX) + (useD ? (size_t)m * sizeof(TU) : 0);
    long epb = 256 >> tpeLog;
    const size_t budget = 48 * 1024;
    if (fixed + ((256 >> tpeLog) < 8 ? (256 >> tpeLog) : 8) * per + 32 > budget) {
        const long e2 = 256 >> tpeLog;
        const long b2 = std::min<long>((N + e2 - 1) / e2, 1L << 16);
        hipLaunchKernelGGL((k_outputs_direct<TX, TU>), dim3((unsigned)std::max<long>(b2, 1)),
                           dim3(256), 0, s, n, m, p, na, oX, zC, zD, N, tpeLog, C, D, X, U, sO,
                           oO, Y);
        check_launch("outputs_direct");
        return;
    }
    if (fixed + epb * per + 32 > budget) epb = (long)((budget - fixed - 32) / per);
    const size_t lds = fixed + (size_t)epb * per + 32;   // + alignment of the parts
    const long blocks = std::min<long>((N + epb - 1) / epb, 1L << 16);
    hipLaunchKernelGGL((k_outputs<TX, TU>), dim3((unsigned)std::max<long>(blocks, 1)), dim3(256),
                       lds, s, n, m, p, na, oX, zC, zD, N, tpeLog, (int)epb, C, D, X, U, sO, oO, Y);
    check_launch("outputs");
}

template void outputs<float, float>(hipStream_t, int, int, int, int, int, bool, bool, long,
                                    const float*, const float*, const float*, const float*, long,
                                    long, float*);
template void outputs<float, double>(hipStream_t, int, int, int, int, int, bool, bool, long,
                                     const float*, const float*, const float*, const double*,
                                     long, long, float*);
template void outputs<double, double>(hipStream_t, int, int, int, int, int, bool, bool, long,
                                      const float*, const float*, const double*, const double*,
                                      long, long, float*);
template void outputs<double, float>(hipStream_t, int, int, int, int, int, bool, bool, long,
                                     const float*, const float*, const double*, const float*,
                                     long, long, float*);

// ------------------------------------------------------------ ABM start-up
// Three classic RK4 steps (AdamsMoulton.cpp:96-103).  mirror: float state and
// float derivative sums (AdamsMoulton.cpp:15-27), stage combinations formed
// in double and stored to float (odeint runge_kutta4<vector<float>>).
// When A fits (pitch n + 1, no bank conflicts), it is staged in LDS once: the
// twelve derivative evaluations are chains of n dependent products per row,
// which from global memory wait on a load each.
template <typename TU>
__device__ __forceinline__ void d_abm_startup(const VBlock vb, bool mirror, int n, int m, double h, bool ldsA,
                              const float* __restrict__ A, const float* __restrict__ B,
                              const float* __restrict__ x0, const TU* __restrict__ Uh,
                              double* __restrict__ X0)
{
#pragma clang fp contract(off)
    extern __shared__ double sh[];
    double* x = sh;            // n
    double* xt = sh + n;       // n
    double* kb = sh + 2 * n;   // 4n: k1..k4
    float* As = reinterpret_cast<float*>(sh + 6 * n);   // n x (n + 1) when ldsA
    const double dh = h / 2.0, dt6 = h / 6.0, dt3 = h / 3.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        x[r] = (double)x0[r];
        X0[r] = (double)x0[r];
    }
    if (ldsA)
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            const int r = t / n;
            As[t + r] = A[t];   // row r at r * (n + 1)
        }
    __syncthreads();
    const float* Ar = ldsA ? As : A;
    const long ap = ldsA ? n + 1 : n;
    auto deriv = [&](const double* xs, int urow, double* out) {
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            if (mirror) {
                float f = 0.0f;
                for (int c = 0; c < n; ++c) {
                    float pr = Ar[r * ap + c] * (float)xs[c];
                    f = f + pr;
                }
                for (int c = 0; c < m; ++c) {
                    float pr = B[(long)r * m + c] * (float)Uh[(long)urow * m + c];
                    f = f + pr;
                }
                out[r] = (double)f;
            } else {
                double f = 0.0;
                for (int c = 0; c < n; ++c) f = fma((double)Ar[r * ap + c], xs[c], f);
                for (int c = 0; c < m; ++c)
                    f = fma((double)B[(long)r * m + c], (double)Uh[(long)urow * m + c], f);
                out[r] = f;
            }
        }
        __syncthreads();
    };
    auto stage = [&](const double* kk, double coef) {
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double v = 1.0 * x[r] + coef * kk[r];
            xt[r] = mirror ? (double)(float)v : v;
        }
        __syncthreads();
    };
    for (int s = 0; s < 3; ++s) {
        double* k1 = kb;
        double* k2 = kb + n;
        double* k3 = kb + 2 * n;
        double* k4 = kb + 3 * n;
        deriv(x, 2 * s, k1);
        stage(k1, dh);
        deriv(xt, 2 * s + 1, k2);
        stage(k2, dh);
        deriv(xt, 2 * s + 1, k3);
        stage(k3, h);
        deriv(xt, 2 * s + 2, k4);
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double v;
            if (mirror)
                v = (double)(float)((((1.0 * x[r] + dt6 * k1[r]) + dt3 * k2[r]) + dt3 * k3[r]) +
                                    dt6 * k4[r]);
            else
                v = x[r] + (h / 6.0) * (k1[r] + 2.0 * k2[r] + 2.0 * k3[r] + k4[r]);
            x[r] = v;
            X0[(long)(s + 1) * n + r] = v;
        }
        __syncthreads();
    }
}

template <typename TU>
__global__ void k_abm_startup(bool mirror, int n, int m, double h, bool ldsA,
                              const float* __restrict__ A, const float* __restrict__ B,
                              const float* __restrict__ x0, const TU* __restrict__ Uh,
                              double* __restrict__ X0)
{
    d_abm_startup<TU>(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, mirror, n, m, h, ldsA, A, B, x0, Uh, X0);
}

template <typename TU>
void abm_startup(hipStream_t s, bool mirror, int n, int m, double h, const float* A,
                 const float* B, const float* x0, const TU* Uh, double* X0)
{
    size_t lds = (size_t)6 * n * sizeof(double);
    const size_t ldsA = (size_t)n * (n + 1) * sizeof(float);
    const bool useA = lds + ldsA <= 65536;
    if (useA) lds += ldsA;
    hipLaunchKernelGGL((k_abm_startup<TU>), dim3(1), dim3(256), lds, s, mirror, n, m, h, useA, A,
                       B, x0, Uh, X0);
    check_launch("abm_startup");
}

template void abm_startup<float>(hipStream_t, bool, int, int, double, const float*, const float*,
                                 const float*, const float*, double*);
template void abm_startup<double>(hipStream_t, bool, int, int, double, const float*, const float*,
                                  const float*, const double*, double*);

template <typename T>
__global__ void k_to_f64(long count, const T* __restrict__ in, double* __restrict__ out)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

template <typename T>
void to_f64(hipStream_t s, long count, const T* in, double* out)
{
    hipLaunchKernelGGL((k_to_f64<T>), dim3(blocks_for(count, 256, 4096)), dim3(256), 0, s, count, in,
                       out);
    check_launch("to_f64");
}

template void to_f64<float>(hipStream_t, long, const float*, double*);

// ----------------------------------------------------------- input samples
__device__ inline uint64_t mix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Samples come in runs of kSampleRun steps of one input, each run starting at
// an absolute step that is a multiple of kSampleRun: one sincos at the run's
// start, then sin g and cos g advance by the rotation through k4 dt per step
// (a few ulps of drift over a run).  A sample's bits depend only on its
// (step, input), not on the grid or on s0.
constexpr int kSampleRun = 16;

__device__ __forceinline__ void d_sample_family(const VBlock vb, int m, const double* __restrict__ coef, double t0, double dt,
                                long s0, long S, double sigma, uint64_t seed,
                                double* __restrict__ out)
{
    const long q0 = s0 / kSampleRun, items = ((s0 + S - 1) / kSampleRun - q0 + 1) * m;
    const long stride = (long)vb.gx * blockDim.x;
    for (long w = (long)vb.x * blockDim.x + threadIdx.x; w < items; w += stride) {
        const long qi = w / m;
        const int j = (int)(w - qi * m);
        const double* k = coef + 6 * j;
        long a = (q0 + qi) * kSampleRun;
        const long end = a + kSampleRun < s0 + S ? a + kSampleRun : s0 + S;
        double sg, cg, sd, cd;
        sincos(fma(k[4], t0 + (double)a * dt, k[5]), &sg, &cg);   // one argument reduction for both
        sincos(k[4] * dt, &sd, &cd);
        for (; a < end; ++a) {
            if (a >= s0) {
                const double t = t0 + (double)a * dt;
                double u = fma(k[3], t, fma(k[1], cg, fma(k[0], sg, k[2])));
                if (sigma != 0.0) {
                    // Box-Muller on two counter-keyed splitmix64 draws
                    uint64_t st = seed * 0xD1B54A32D192ED03ull + (uint64_t)a * 0x9E3779B97F4A7C15ull +
                                  (uint64_t)j * 0xA24BAED4963EE407ull;
                    const double u1 = ((double)(mix64(st) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
                    const double u2 = (double)(mix64(st) >> 11) * (1.0 / 9007199254740992.0);
                    u += sigma * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
                }
                out[(a - s0) * m + j] = u;
            }
            const double s1 = fma(sg, cd, cg * sd), c1 = fma(cg, cd, -(sg * sd));
            sg = s1;
            cg = c1;
        }
    }
}

// work items (runs of one input) of S samples per input starting at step s0
static inline long sample_items(int m, long s0, long S)
{
    return ((s0 + S - 1) / kSampleRun - s0 / kSampleRun + 1) * m;
}

__global__ void k_sample_family(int m, const double* __restrict__ coef, double t0, double dt,
                                long s0, long S, double sigma, uint64_t seed,
                                double* __restrict__ out)
{
    d_sample_family(VBlock{(long)blockIdx.x, (long)blockIdx.y, (long)gridDim.x}, m, coef, t0, dt, s0, S, sigma, seed, out);
}

void sample_family(hipStream_t s, int m, const double* coef, double t0, double dt, long s0,
                   long S, double sigma, uint64_t seed, double* out)
{
    if (S * m <= 0) return;
    hipLaunchKernelGGL(k_sample_family, dim3(blocks_for(sample_items(m, s0, S), 256, 8192)), dim3(256), 0, s, m,
                       coef, t0, dt, s0, S, sigma, seed, out);
    check_launch("sample_family");
}


// ====================================================== grouped launches
// One grid for many systems.  System s of a launch owns the blocks
// [start[s], start[s + 1]) and the argument record args[s]; a work-group
// finds its system by a binary search over start (the same for every lane,
// so it runs on scalar loads) and runs the same body as a single-system
// launch, with its block's place in that system's grid.  The kernels a
// system runs, and how its work is split into blocks, depend only on its own
// shape, so its results do not depend on the other systems of the launch.
namespace g {

namespace {

__device__ __forceinline__ int system_of(const unsigned* __restrict__ start, int nsys)
{
    const unsigned b = blockIdx.x;
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ VBlock block_of(const unsigned* __restrict__ start, int s, long gx)
{
    const long lb = (long)blockIdx.x - (long)start[s];
    return VBlock{lb % gx, lb / gx, gx};
}

// Argument records (one per system of a launch; gx: the system's grid width).
struct Prep {        // kind 0: out = (double)in; 1: out = v; 2: out[r][c] = in[c][r], c < cols, padded to pp
    int kind, rows, cols, pp;
    long count;
    double v;
    const float* in;
    double* out;
    long gx;
};
struct Gemm {
    int r, kk, c;
    const double *X, *Y;
    double* Z;
    GemmLd ld;
    long gx;
};
struct Rk4Comb {
    int n, m;
    Rk4Coef k;
    const double *A, *A2, *A3, *A4, *B, *AB, *A2B, *A3B;
    double *MT, *Bbar;
    long gx;
};
struct AbmComb {
    int n, m;
    AbmCoef c;
    const double *A, *A2, *B, *AB;
    double *MT, *Bbar;
    long gx;
};
struct Transpose {
    int rows, cols;
    const double* in;
    double* out;
    long gx;
};
struct Powers {
    int na, L;
    double* PT;
    long gx;
};
struct Tables {
    int na, nap, nrp, nlev, deltaBelow;
    const double* PT64;
    float *PT32, *PTm;
    long gx;
};
struct Sample {
    int m;
    const double* coef;
    double t0, dt;
    long S;
    double sigma;
    uint64_t seed;
    double* out;
    long gx;
};
struct Startup {
    int n, m;
    double h;
    bool ldsA;
    const float *A, *B, *x0;
    const double* Uh;
    double* X0;
    long gx;
};
struct StartOut {
    int n, m, p, pp;
    bool useC, useD;
    const double *CT, *DT, *X0, *U;
    float* Y;
    long gx;
};
struct Leaves {
    int n, m, kq, na, off;
    long nReal, nPad;
    int groups, pitch;
    unsigned magic;
    int kg;
    LeafComb G;
    bool writeZeros;
    int zlo;
    const double *BbT, *U;
    long sL, oL, rs;
    float* X;
    long gx;
};
struct Chunks {      // ABM: n, A64, A264; dense: na (in n), P64 (in A64)
    int n, lc;
    long nReal;
    const double *A64, *A264;
    AbmFactors f;
    float* X;
    long gx;
};
struct Level {
    int na, ldp;
    long d, np;
    const void* PT;
    float *S, *X, *Sn;
    long npn;
    long gx;
};
struct Gather {
    int na;
    long d, np;
    const float* X;
    float* S;
    long gx;
};
struct Top {
    int na, l0, L, l64;
    long nReal;
    const float* PT32;
    const double* PT64;
    float* X;
    const double* seed;
    double* tot;
    int CC;
    long gx;
};
struct Root {
    int na;
    long last;
    const double* seed;
    float* X;
    double* tot;
    long gx;
};
struct Outputs {
    int n, m, p, pp, na, oX;
    bool useC, useD;
    long N;
    const double *CT, *DT;
    const float* X;
    const double* U;
    long sO, oO;
    float* Y;
    long gx;
};

#define PARODE_GROUP(T)                          \
    const int sys_ = system_of(st, ns);          \
    const T& a = args[sys_];                     \
    const VBlock vb = block_of(st, sys_, a.gx)

__global__ __launch_bounds__(256) void kg_prep(const Prep* __restrict__ args,
                                               const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Prep);
    const long total = a.kind == 2 ? (long)a.rows * a.pp : a.count;
    for (long i = vb.x * 256 + threadIdx.x; i < total; i += vb.gx * 256) {
        double v;
        if (a.kind == 0) {
            v = (double)a.in[i];
        } else if (a.kind == 1) {
            v = a.v;
        } else {
            const long r = i / a.pp, c = i - r * a.pp;
            v = c < a.cols ? (double)a.in[c * a.rows + r] : 0.0;
        }
        a.out[i] = v;
    }
}

// FP64 products of one phase in one launch: each system's 64 x 64 tiles, or,
// for a narrow product, its K-split 16 x 16 tiles (the two tilings in separate
// launches would run one after the other on the stream)
__global__ __launch_bounds__(256) void kg_gemm(const Gemm* __restrict__ args,
                                               const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Gemm);
    if (gemm_wide(a.r, a.kk, a.c))
        d_gemm_f64_t64(vb, a.r, a.kk, a.c, a.X, a.Y, a.Z, a.ld);
    else
        d_gemm_f64_mfma(vb, a.r, a.kk, a.c, a.X, a.Y, a.Z, a.ld);
}

__global__ __launch_bounds__(512) void kg_gemm_ks2(const Gemm* __restrict__ args,
                                                   const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Gemm);
    d_gemm_f64_ks<2>(vb, a.r, a.kk, a.c, a.X, a.Y, a.Z, a.ld);
}

__global__ __launch_bounds__(256) void kg_rk4_combine(const Rk4Comb* __restrict__ args,
                                                      const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Rk4Comb);
    d_rk4_combine<double>(vb, false, a.n, a.m, a.k, a.A, a.A2, a.A3, a.A4, a.B, a.AB, a.A2B,
                          a.A3B, a.MT, a.Bbar, nullptr);
}

__global__ __launch_bounds__(256) void kg_abm_combine(const AbmComb* __restrict__ args,
                                                      const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(AbmComb);
    d_abm_combine<double>(vb, false, a.n, a.m, a.c, a.A, a.A2, a.B, a.AB, a.MT, a.Bbar, nullptr);
}

__global__ __launch_bounds__(256) void kg_transpose(const Transpose* __restrict__ args,
                                                    const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Transpose);
    d_transpose<double>(vb, a.rows, a.cols, a.in, a.out);
}

__global__ __launch_bounds__(1024) void kg_powers_small(const Powers* __restrict__ args,
                                                        const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Powers);
    d_powers_small(vb, a.na, a.L, a.PT);
}

__global__ __launch_bounds__(256) void kg_tables(const Tables* __restrict__ args,
                                                 const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Tables);
    d_level_tables(vb, a.na, a.nap, a.nrp, a.nlev, a.deltaBelow, a.PT64, a.PT32, a.PTm);
}

__global__ __launch_bounds__(256) void kg_sample(const Sample* __restrict__ args,
                                                 const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Sample);
    d_sample_family(vb, a.m, a.coef, a.t0, a.dt, 0, a.S, a.sigma, a.seed, a.out);
}

__global__ __launch_bounds__(256) void kg_startup(const Startup* __restrict__ args,
                                                  const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Startup);
    d_abm_startup<double>(vb, false, a.n, a.m, a.h, a.ldsA, a.A, a.B, a.x0, a.Uh, a.X0);
}

// The outputs of the ABM start-up states x_0, x_1, x_2 (the first n-blocks of
// X0), whose inputs are rows 0, 2, 4 of the h/2-grid samples
// (AdamsMoulton.cpp:106-116); FP64 sums, C terms then D terms.
__global__ __launch_bounds__(256) void kg_start_outputs(const StartOut* __restrict__ args,
                                                        const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(StartOut);
    (void)vb;
    for (int t = threadIdx.x; t < 3 * a.p; t += 256) {
        const int i = t / a.p, j = t - i * a.p;
        double v = 0.0;
        if (a.useC)
            for (int c = 0; c < a.n; ++c)
                v = fma(a.CT[(long)c * a.pp + j], a.X0[(long)i * a.n + c], v);
        if (a.useD)
            for (int c = 0; c < a.m; ++c)
                v = fma(a.DT[(long)c * a.pp + j], a.U[(long)(2 * i) * a.m + c], v);
        a.Y[(long)i * a.p + j] = (float)v;
    }
}

template <bool FAC>
__global__ __launch_bounds__(256, 3) void kg_leaves(const Leaves* __restrict__ args,
                                                    const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Leaves);
    d_leaves_mfma<float, double, FAC>(vb, a.n, a.m, a.kq, a.na, a.off, a.nReal, a.nPad, a.groups,
                                      a.pitch, a.magic, a.kg, a.G, a.writeZeros, a.zlo, a.BbT,
                                      a.U, a.sL, a.oL, a.rs, a.X);
}

template <int JT>
__global__ __launch_bounds__(256) void kg_outputs_gemm(const Outputs* __restrict__ args,
                                                       const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Outputs);
    d_outputs_gemm<float, double, JT>(vb, a.n, a.m, a.p, a.pp, a.na, a.oX, a.useC, a.useD, a.N,
                                      a.CT, a.DT, a.X, a.U, a.sO, a.oO, a.Y);
}

__global__ __launch_bounds__(256) void kg_leaves_simt(const Leaves* __restrict__ args,
                                                      const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Leaves);
    // (Leaves::pitch holds the thread groups' width, log 2)
    d_leaves<float, double, double>(vb, a.n, a.m, a.kq, a.na, a.off, a.nReal, a.nPad, a.pitch,
                                    a.groups, a.writeZeros, a.zlo, a.BbT, a.U, a.sL, a.oL, a.rs,
                                    a.X);
}

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 4) void kg_chunks_fac(const Chunks* __restrict__ args,
                                                        const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Chunks);
    d_abm_chunks_fac<NP, PASS2>(vb, a.n, a.lc, a.nReal, a.A64, a.A264, a.f, a.X);
}

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 2) void kg_chunks_two(const Chunks* __restrict__ args,
                                                        const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Chunks);
    d_abm_chunks_two<NP, PASS2>(vb, a.n, a.lc, a.nReal, a.A64, a.f, a.X);
}

template <int NP, bool PASS2>
__global__ __launch_bounds__(NP * 4) void kg_chunks_dense(const Chunks* __restrict__ args,
                                                          const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Chunks);
    d_chunks_dense<NP, PASS2>(vb, a.n, a.lc, 1, a.nReal, a.A64, a.X);
}

// FP32 levels on 16 x 16 tiles of the padded MFMA table (TT = float), FP64
// levels on 16 x 16 tiles of the transposed FP64 power (TT = double).
template <typename TT, bool DOWN, bool DELTA, bool SMALL>
__global__ __launch_bounds__(256) void kg_level(const Level* __restrict__ args,
                                                const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Level);
    d_level_ks<TT, float, DOWN, DELTA, SMALL>(vb, a.na, a.ldp, a.d, a.np,
                                              static_cast<const TT*>(a.PT), a.S, a.X,
                                              DOWN ? a.Sn : nullptr, DOWN ? a.npn : 0L);
}

template <bool DOWN, bool DELTA>
__global__ __launch_bounds__(256) void kg_level_gemm(const Level* __restrict__ args,
                                                     const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Level);
    d_level_gemm<DOWN, DELTA>(vb, a.na, a.ldp, a.d, a.np, static_cast<const float*>(a.PT), a.S,
                              a.X, DOWN ? a.Sn : nullptr, DOWN ? a.npn : 0L);
}

__global__ __launch_bounds__(256) void kg_gather(const Gather* __restrict__ args,
                                                 const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Gather);
    d_gather_rt<float>(vb, a.na, a.d, a.np, a.X, a.S);
}

template <bool MFMA>
__global__ __launch_bounds__(kTopThreads) void kg_top(const Top* __restrict__ args,
                                                      const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Top);
    d_top<float, MFMA>(vb, a.na, a.l0, a.L, a.l64, a.nReal, a.PT32, a.PT64, a.X, a.seed, a.tot,
                       a.CC);
}

__global__ __launch_bounds__(256) void kg_root(const Root* __restrict__ args,
                                               const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Root);
    d_root<float, double>(vb, a.na, a.last, a.seed, a.X, a.tot);
}

__global__ __launch_bounds__(kOutLanes) void kg_out_lanes(const Outputs* __restrict__ args,
                                                          const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Outputs);
    d_outputs_lanes<float, double>(vb, a.n, a.m, a.p, a.pp, a.na, a.oX, a.useC, a.useD, a.N, a.CT,
                                   a.DT, a.X, a.U, a.sO, a.oO, a.Y);
}

__global__ __launch_bounds__(256) void kg_out_mfma(const Outputs* __restrict__ args,
                                                   const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Outputs);
    d_outputs_mfma(vb, a.n, a.m, a.p, a.pp, a.na, a.oX, a.useC, a.useD, a.N, a.CT, a.DT, a.X,
                   a.U, a.sO, a.oO, a.Y);
}

template <int JG>
__global__ __launch_bounds__(kOutLanes) void kg_out_kc(const Outputs* __restrict__ args,
                                                       const unsigned* __restrict__ st, int ns)
{
    PARODE_GROUP(Outputs);
    d_outputs_kc<float, double, JG>(vb, a.n, a.m, a.p, a.pp, a.na, a.oX, a.useC, a.useD, a.N,
                                    a.CT, a.DT, a.X, a.U, a.sO, a.oO, a.Y);
}

#undef PARODE_GROUP

// The record of the launch (kind, flag) in phase ph, created on first use.
template <typename A>
Launch& slot(Phase& ph, int kind, bool flag, int cls)
{
    Launch& L = ph[{kind, (flag ? 1 : 0) + 2 * cls}];
    if (L.kind < 0) {
        L.kind = kind;
        L.flag = flag;
        L.argSize = sizeof(A);
    }
    if (L.argSize != sizeof(A)) throw std::logic_error("grouped launch: argument type mismatch");
    return L;
}

// Systems whose dynamic LDS differs go to separate launches by size class
// (up to 8, 16, 32, 48, 64 KiB ...): a launch reserves its largest system's
// LDS for every work-group, which would cap the occupancy of the others.
int lds_class(size_t lds)
{
    if (lds == 0) return 0;
    const size_t kib = (lds + 1023) / 1024;
    return kib <= 8 ? 1 : kib <= 16 ? 2 : kib <= 32 ? 3 : kib <= 48 ? 4 : 5 + (int)((kib - 49) / 16);
}

template <typename A>
void append(Phase& ph, int kind, bool flag, const A& a, long blocks, size_t lds = 0,
            float cost = 0.0f)
{
    if (blocks < 1) return;
    // (the top of the tree runs one 1024-thread work-group per system, one per
    // CU whatever its LDS: its systems share one launch)
    const bool oneClass = kind == kTopMfma || kind == kTopSimt;
    Launch& L = slot<A>(ph, kind, flag, oneClass ? 0 : lds_class(lds));
    const size_t o = L.args.size();
    L.args.resize(o + sizeof(A));
    std::memcpy(L.args.data() + o, &a, sizeof(A));
    const unsigned long next = (unsigned long)L.start.back() + (unsigned long)blocks;
    if (next > 0x7fffffffUL) throw std::length_error("grouped launch: too many blocks");
    L.start.push_back((unsigned)next);
    L.cost.push_back(cost);
    L.lds = std::max(L.lds, lds);
}

}  // namespace

void pack_by_cost(const Launch& L, unsigned char* args, unsigned* start)
{
    const int ns = L.nsys();
    std::vector<int> idx(ns);
    for (int i = 0; i < ns; ++i) idx[i] = i;
    if ((int)L.cost.size() == ns)
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return L.cost[a] > L.cost[b]; });
    start[0] = 0;
    for (int k = 0; k < ns; ++k) {
        const int s = idx[k];
        std::memcpy(args + (size_t)k * L.argSize, L.args.data() + (size_t)s * L.argSize,
                    L.argSize);
        start[k + 1] = start[k] + (L.start[s + 1] - L.start[s]);
    }
}

void merge_launch(Launch& into, Launch&& from)
{
    if (from.nsys() <= 0) return;
    if (into.kind < 0) {
        into = std::move(from);
        return;
    }
    if (into.kind != from.kind || into.flag != from.flag || into.argSize != from.argSize)
        throw std::logic_error("grouped launch: merging different launches");
    const unsigned base = into.blocks();
    if ((unsigned long)base + from.blocks() > 0x7fffffffUL)
        throw std::length_error("grouped launch: too many blocks");
    into.args.insert(into.args.end(), from.args.begin(), from.args.end());
    for (int k = 1; k <= from.nsys(); ++k) into.start.push_back(base + from.start[k]);
    into.cost.insert(into.cost.end(), from.cost.begin(), from.cost.end());
    into.lds = std::max(into.lds, from.lds);
}

void add_cast(Phase& ph, long count, const float* in, double* out)
{
    if (count <= 0) return;
    const long b = blocks_for(count, 256, 512);
    append(ph, kPrep, false, Prep{0, 0, 0, 0, count, 0.0, in, out, b}, b);
}

void add_fill(Phase& ph, long count, double v, double* out)
{
    if (count <= 0) return;
    const long b = blocks_for(count, 256, 512);
    append(ph, kPrep, false, Prep{1, 0, 0, 0, count, v, nullptr, out, b}, b);
}

void add_transpose_pad(Phase& ph, int rows, int cols, int pp, const float* in, double* out)
{
    if (rows <= 0 || pp <= 0) return;
    const long b = blocks_for((long)rows * pp, 256, 1024);
    append(ph, kPrep, false, Prep{2, rows, cols, pp, 0, 0.0, in, out, b}, b);
}

void add_gemm(Phase& ph, int r, int kk, int c, const double* X, const double* Y, double* Z)
{
    add_gemm_ld(ph, r, kk, c, X, kk, Y, c, Z, c, nullptr, 0);
}

void add_gemm_ld(Phase& ph, int r, int kk, int c, const double* X, int ldx, const double* Y,
                 int ldy, double* Z, int ldz, const double* add, int shift)
{
    if (r <= 0 || c <= 0) return;
    const GemmLd ld{ldx, ldy, ldz, add, shift};
    // (a work-group's matrix-core instructions per wave: kk for a 64 x 64 tile,
    // kk / 16 for a K-split 16 x 16 tile)
    if (gemm_wide(r, kk, c)) {
        const long gx = (c + kG64 - 1) / kG64, gy = (r + kG64 - 1) / kG64;
        append(ph, kGemm, false, Gemm{r, kk, c, X, Y, Z, ld, gx}, gx * gy, 0, (float)kk);
    } else {
        const long gx = (c + 15) / 16, gy = (r + 15) / 16;
        append(ph, kGemm, false, Gemm{r, kk, c, X, Y, Z, ld, gx}, gx * gy, 0, kk / 16.0f);
    }
}

void add_rk4_combine(Phase& ph, int n, int m, double h, const double* A, const double* A2,
                     const double* A3, const double* A4, const double* B, const double* AB,
                     const double* A2B, const double* A3B, double* MT, double* Bbar)
{
    const long b = blocks_for((long)n * n + (long)n * m, 256);
    append(ph, kRk4Combine, false,
           Rk4Comb{n, m, rk4_coef(h), A, A2, A3, A4, B, AB, A2B, A3B, MT, Bbar, b}, b);
}

void add_abm_combine(Phase& ph, int n, int m, double h, const double* A, const double* A2,
                     const double* B, const double* AB, double* MT, double* Bbar)
{
    const long b = blocks_for((long)n * n + (long)n * m, 256);
    append(ph, kAbmCombine, false, AbmComb{n, m, abm_coef(h), A, A2, B, AB, MT, Bbar, b}, b);
}

void add_transpose(Phase& ph, int rows, int cols, const double* in, double* out)
{
    if ((long)rows * cols <= 0) return;
    const long b = blocks_for((long)rows * cols, 256, 512);
    append(ph, kTranspose, false, Transpose{rows, cols, in, out, b}, b);
}

bool add_powers_small(Phase& ph, int na, int L, double* PT)
{
    if (na > kPowersSmallMax) return false;
    if (L <= 0) return true;
    const int n16 = (na + 15) / 16 * 16;
    append(ph, kPowersSmall, false, Powers{na, L, PT, 1}, 1,
           (size_t)n16 * (n16 + 1) * sizeof(double));
    return true;
}

void add_tables(Phase& ph, int na, int nap, int nrp, int nlev, int deltaBelow, const double* PT64,
                float* PT32, float* PTm, int first)
{
    if (first < 0) first = 0;
    if (nlev - first <= 0) return;
    if (!PTm) nap = nrp = na;
    // the same tables from level `first` on: every pointer moved by `first` levels
    const long nn = (long)na * na, per = (long)nrp * nap;
    PT64 += first * nn;
    if (PT32) PT32 += first * nn;
    if (PTm) PTm += first * per;
    nlev -= first;
    deltaBelow = deltaBelow > first ? deltaBelow - first : 0;
    const long b = blocks_for((long)nrp * nap * nlev, 256, 512);
    append(ph, kTables, false, Tables{na, nap, nrp, nlev, deltaBelow, PT64, PT32, PTm, b}, b);
}

void add_sample(Phase& ph, int m, const double* coef, double t0, double dt, long S, double sigma,
                uint64_t seed, double* out)
{
    if (S * m <= 0) return;
    const long b = blocks_for(sample_items(m, 0, S), 256, 256);
    append(ph, kSample, false, Sample{m, coef, t0, dt, S, sigma, seed, out, b}, b);
}

void add_startup(Phase& ph, int n, int m, double h, const float* A, const float* B,
                 const float* x0, const double* Uh, double* X0)
{
    size_t lds = (size_t)6 * n * sizeof(double);
    const size_t ldsA = (size_t)n * (n + 1) * sizeof(float);
    const bool useA = lds + ldsA <= 65536;
    if (useA) lds += ldsA;
    append(ph, kStartup, false, Startup{n, m, h, useA, A, B, x0, Uh, X0, 1}, 1, lds);
}

void add_start_outputs(Phase& ph, int n, int m, int p, int pp, bool useC, bool useD,
                       const double* CT, const double* DT, const double* X0, const double* U,
                       float* Y)
{
    if (p <= 0) return;
    append(ph, kStartOut, false, StartOut{n, m, p, pp, useC, useD, CT, DT, X0, U, Y, 1}, 1);
}

bool leaves_fit(int n, int m, int kq, int kg, long sL, long rs)
{
    if (m <= 0 || m >= 65536) return false;
    const int rowTiles = (n + 15) / 16;
    int groups = rowTiles >= 4 ? 1 : 4 / rowTiles;
    const int pitch = m | 1;
    while (groups > 1 &&
           leaves_mfma_lds(16L * kLeafTiles * groups, m, kq, kg, sL, rs, pitch) > 65536)
        groups >>= 1;
    const long perBlock = 16L * kLeafTiles * groups;
    const size_t lds = leaves_mfma_lds(perBlock, m, kq, kg, sL, rs, pitch);
    return lds <= 65536 && lds / sizeof(double) < 65536 && perBlock * m < 65536;
}

bool add_leaves(Phase& ph, int n, int m, int kq, int kg, int na, int off, long nReal, long nPad,
                const double* Bf, const LeafComb& G, const double* U, long sL, long oL, long rs,
                float* X, bool writeZeros)
{
    if (m <= 0 || nReal < 1) return false;
    const int rowTiles = (n + 15) / 16;
    int groups = rowTiles >= 4 ? 1 : 4 / rowTiles;
    const int pitch = m | 1;
    while (groups > 1 &&
           leaves_mfma_lds(16L * kLeafTiles * groups, m, kq, kg, sL, rs, pitch) > 65536)
        groups >>= 1;
    const long perBlock = 16L * kLeafTiles * groups;
    const size_t lds = leaves_mfma_lds(perBlock, m, kq, kg, sL, rs, pitch);
    if (lds > 65536 || m >= 65536 || lds / sizeof(double) >= 65536 || perBlock * m >= 65536)
        return false;
    const unsigned magic = (unsigned)(((1ULL << 32) + m - 1) / m);
    // persistent work-groups: a few hundred per system (each walks its blocks,
    // the next block's samples loaded during the current one), so that the
    // per-work-group start-up is paid rarely
    const long nb = (nPad + perBlock - 1) / perBlock;
    const long b = std::max<long>(std::min<long>(nb, 256), 1);
    // without writeZeros the first reader (a chunk pass) reads only rows
    // [off, off + n): no other row is written for real leaves
    append(ph, kg > 0 ? kLeavesFac : kLeavesPlain, false,
           Leaves{n, m, kq, na, off, nReal, nPad, groups, pitch, magic, kg, G, writeZeros,
                  writeZeros ? off / 32 * 32 : off, Bf, U, sL, oL, rs, X, b},
           b, lds, (float)((nb + b - 1) / b));
    return true;
}

// The vector-ALU leaves (k_leaves) from B-bar itself, for input counts whose
// staged samples do not fit the matrix-core kernel: thread groups of 2^tpeLog
// (in Leaves::pitch) lanes, Leaves::groups of them per work-group.
bool add_leaves_simt(Phase& ph, int n, int m, int kq, int na, int off, long nReal, long nPad,
                     const double* BbT, const double* U, long sL, long oL, long rs, float* X,
                     bool writeZeros)
{
    if (m <= 0 || nReal < 1) return false;
    int tpeLog = 0;
    while ((1 << tpeLog) < n && tpeLog < 8) ++tpeLog;
    int groups = 256 >> tpeLog;
    auto lds_for = [&](int g) {
        return (size_t)(((long)g * kLeafEB - 1) * sL + (long)(kq - 1) * rs + 1) * m *
               sizeof(double);
    };
    while (groups > 1 && lds_for(groups) > 65536) groups >>= 1;
    if (lds_for(groups) > 65536) return false;
    const long perBlock = (long)groups * kLeafEB;
    const long nb = (nPad + perBlock - 1) / perBlock;
    const long b = std::max<long>(std::min<long>(nb, 256), 1);
    append(ph, kLeavesSimt, false,
           Leaves{n, m, kq, na, off, nReal, nPad, groups, tpeLog, 0u, 0, LeafComb{}, writeZeros,
                  writeZeros ? off / 32 * 32 : off, BbT, U, sL, oL, rs, X, b},
           b, lds_for(groups), (float)((nb + b - 1) / b));
    return true;
}

int chunks_kind(bool abm, int n)
{
    static const int kNp[] = {16, 32, 48, 64, 96, 128, 160, 192, 224, 256};
    if (abm ? !abm_chunks_factored_supported(n) : !chunks_dense_supported(n)) return -1;
    int t = 0;
    while (t < 9 && kNp[t] < n) ++t;
    return (abm ? kChunksFac16 : kDense16) + t;
}

bool add_abm_chunks(Phase& ph, bool pass2, int n, int lc, long nReal, long nPad, double h,
                    const double* A64, const double* A264, float* X)
{
    if (!abm_chunks_factored_supported(n) || lc < 1 || (nPad >> lc) < 1 || nReal < 1 ||
        nReal > nPad)
        return false;
    // work-groups for the chunks that hold a real element (the others return at once)
    const long b = (((nReal + (1L << lc) - 1) >> lc) + kChunkCols - 1) / kChunkCols;
    const Chunks a{n, lc, nReal, A64, A264, abm_chunk_factors(h), X, b};
    // padded state sizes: factored up to 128, two products above
    static const int kNp[] = {16, 32, 48, 64, 96, 128, 160, 192, 224, 256};
    static const int kKind[] = {kChunksFac16, kChunksFac32, kChunksFac48, kChunksFac64,
                                kChunksFac96, kChunksFac128, kChunksTwo160, kChunksTwo192,
                                kChunksTwo224, kChunksTwo256};
    int t = 0;
    while (kNp[t] < n) ++t;
    const int np = kNp[t];
    const size_t lds = np <= 128
                           ? (size_t)2 * kChunkCols * chunk_pitch(np / 2, 2 * np) * sizeof(float)
                           : (size_t)2 * kChunkCols * chunk_pitch(np / 4, np) * sizeof(float);
    append(ph, kKind[t], pass2, a, b, lds, (float)(1L << lc));
    return true;
}

bool add_dense_chunks(Phase& ph, bool pass2, int na, int lc, long nReal, long nPad,
                      const double* P64, float* X)
{
    if (!chunks_dense_supported(na) || lc < 1 || (nPad >> lc) < 1 || nReal < 1 || nReal > nPad)
        return false;
    const long b = (((nReal + (1L << lc) - 1) >> lc) + kChunkCols - 1) / kChunkCols;
    static const int kNp[] = {16, 32, 48, 64, 96, 128, 160, 192, 224, 256};
    static const int kKind[] = {kDense16, kDense32, kDense48, kDense64, kDense96, kDense128,
                                kDense160, kDense192, kDense224, kDense256};
    int t = 0;
    while (kNp[t] < na) ++t;
    append(ph, kKind[t], pass2, Chunks{na, lc, nReal, P64, nullptr, AbmFactors{}, X, b}, b,
           (size_t)kChunkCols * chunk_pitch(kNp[t] / 4, kNp[t]) * sizeof(float),
           (float)(1L << lc));
    return true;
}

// pairs from which a grouped FP32 level takes the LDS-staged 64 x 64 tiles
constexpr long kLevelGemmPairs = 256;

void add_level(Phase& ph, bool down, bool f64, bool delta, int na, int ldp, long d, long np,
               const void* PT, float* S, float* X, float* Sn, long npn)
{
    if (np <= 0) return;
    if (!f64 && ldp % 32 != 0) throw std::invalid_argument("add_level: FP32 table must be padded");
    if (down && (!S || (Sn && (Sn == S || npn < 0 || npn > 2 * np))))
        throw std::invalid_argument("add_level: bad copy buffers");
    // FP32 with many pairs: 64 x 64 tiles staged in LDS
    if (!f64 && np >= kLevelGemmPairs && na >= 64) {
        const long gx = (np + kLvG - 1) / kLvG, gy = (na + kLvG - 1) / kLvG;
        append(ph, delta ? kLevelGDelta : kLevelG, down,
               Level{na, ldp, d, np, PT, S, X, Sn, Sn ? npn : 0, gx}, gx * gy);
        return;
    }
    // FP32: 32 x 32 tiles (half the table and state traffic per product) once
    // the system's own grid of them has 64 work-groups, else 16 x 16
    const bool wide = !f64 && ((np + 31) / 32) * ((na + 31) / 32) >= 64;
    const int T = wide ? 32 : 16;
    const long gx = (np + T - 1) / T, gy = (na + T - 1) / T;
    const int kind = f64 ? kLevel64 : delta ? (wide ? kLevel32DeltaW : kLevel32Delta)
                                            : (wide ? kLevel32W : kLevel32);
    append(ph, kind, down, Level{na, ldp, d, np, PT, S, X, Sn, Sn ? npn : 0, gx}, gx * gy);
}

void add_gather(Phase& ph, int na, long d, long np, const float* X, float* S)
{
    if (np <= 0) return;
    const long b = blocks_for(np * na, 256, 1024);
    append(ph, kGather, false, Gather{na, d, np, X, S, b}, b);
}

void add_top(Phase& ph, int na, int l0, int L, int l64, long nReal, const float* PT32,
             const double* PT64, float* X, const double* seed, double* tot)
{
    const int rowThreads = na < kTopThreads ? na : kTopThreads;
    const int groups = kTopThreads / rowThreads;
    const int SP = na >= 16 ? na + 1 : na;
    int CC = na >= 16 ? top_mfma_cols(na) : std::min(groups * kTopMaxCols, 6144 / na);
    if (CC < 1) CC = 1;
    append(ph, na >= 16 ? kTopMfma : kTopSimt, false,
           Top{na, l0, L, l64, nReal, PT32, PT64, X, seed, tot, CC, 1}, 1,
           (size_t)CC * SP * sizeof(double));
}

void add_root(Phase& ph, int na, long last, const double* seed, float* X, double* tot)
{
    append(ph, kRoot, false, Root{na, last, seed, X, tot, 1}, 1);
}

bool add_outputs(Phase& ph, int n, int m, int p, int pp, int na, int oX, bool useC, bool useD,
                 long N, const double* CT, const double* DT, const float* X, const double* U,
                 long sO, long oO, float* Y)
{
    if (N <= 0) return true;
    if (p <= 0 || pp % 8 != 0 || pp < p) return false;
    const long b = std::max<long>(std::min<long>((N + kOutLanes - 1) / kOutLanes, 1024), 1);
    const Outputs a{n, m, p, pp, na, oX, useC, useD, N, CT, DT, X, U, sO, oO, Y, b};
    if (useC && n >= 32) {   // wide states: 64-element tiles on the FP64 matrix cores
        const int jt = outputs_gemm_width(p);
        Outputs ag = a;
        ag.gx = (p + 16 * jt - 1) / (16 * jt);
        append(ph, jt == 1 ? kOutputsGemm1 : jt == 2 ? kOutputsGemm2 : kOutputsGemm4, false, ag,
               ag.gx * ((N + kOG - 1) / kOG));
    } else if (n >= 32) {   // wide states without C: the FP64 matrix cores
        const long bm = std::max<long>(std::min<long>((N + 63) / 64, 512), 1);
        Outputs am = a;
        am.gx = bm;
        append(ph, kOutputsMfma, false, am, bm);
    } else if (n <= kOutLanes) {
        append(ph, kOutputsLanes, false, a, b, (size_t)kOutLanes * (n + 1) * sizeof(float));
    } else {
        const int kind = pp % 32 == 0 ? kOutputsKc32 : pp % 24 == 0 ? kOutputsKc24
                       : pp % 16 == 0 ? kOutputsKc16 : kOutputsKc8;
        append(ph, kind, false, a, b);
    }
    return true;
}

void run(hipStream_t s, const Launch& L, const void* dArgs, const unsigned* dStart)
{
    const int ns = L.nsys();
    if (ns <= 0 || L.blocks() == 0) return;
    const dim3 grid(L.blocks());
    const size_t lds = L.lds;
    const bool f = L.flag;
#define PARODE_RUN(KERNEL, T, THREADS)                                                        \
    hipLaunchKernelGGL((KERNEL), grid, dim3(THREADS), lds, s, static_cast<const T*>(dArgs),   \
                       dStart, ns)
#define PARODE_RUN2(K0, K1, T, THREADS)      \
    do {                                     \
        if (f)                               \
            PARODE_RUN(K1, T, THREADS);      \
        else                                 \
            PARODE_RUN(K0, T, THREADS);      \
    } while (0)
    switch (L.kind) {
    case kPrep: PARODE_RUN(kg_prep, Prep, 256); break;
    case kGemm: PARODE_RUN(kg_gemm, Gemm, 256); break;
    case kGemmKs2: PARODE_RUN(kg_gemm_ks2, Gemm, 512); break;
    case kRk4Combine: PARODE_RUN(kg_rk4_combine, Rk4Comb, 256); break;
    case kAbmCombine: PARODE_RUN(kg_abm_combine, AbmComb, 256); break;
    case kTranspose: PARODE_RUN(kg_transpose, Transpose, 256); break;
    case kPowersSmall: PARODE_RUN(kg_powers_small, Powers, 1024); break;
    case kTables: PARODE_RUN(kg_tables, Tables, 256); break;
    case kSample: PARODE_RUN(kg_sample, Sample, 256); break;
    case kStartup: PARODE_RUN(kg_startup, Startup, 256); break;
    case kStartOut: PARODE_RUN(kg_start_outputs, StartOut, 256); break;
    case kLeavesPlain: PARODE_RUN(kg_leaves<false>, Leaves, 256); break;
    case kLeavesFac: PARODE_RUN(kg_leaves<true>, Leaves, 256); break;
    case kLeavesSimt: PARODE_RUN(kg_leaves_simt, Leaves, 256); break;
    case kOutputsGemm1: PARODE_RUN(kg_outputs_gemm<1>, Outputs, 256); break;
    case kOutputsGemm2: PARODE_RUN(kg_outputs_gemm<2>, Outputs, 256); break;
    case kOutputsGemm4: PARODE_RUN(kg_outputs_gemm<4>, Outputs, 256); break;
    case kChunksFac16: PARODE_RUN2((kg_chunks_fac<16, false>), (kg_chunks_fac<16, true>), Chunks, 64); break;
    case kChunksFac32: PARODE_RUN2((kg_chunks_fac<32, false>), (kg_chunks_fac<32, true>), Chunks, 128); break;
    case kChunksFac48: PARODE_RUN2((kg_chunks_fac<48, false>), (kg_chunks_fac<48, true>), Chunks, 192); break;
    case kChunksFac64: PARODE_RUN2((kg_chunks_fac<64, false>), (kg_chunks_fac<64, true>), Chunks, 256); break;
    case kChunksFac96: PARODE_RUN2((kg_chunks_fac<96, false>), (kg_chunks_fac<96, true>), Chunks, 384); break;
    case kChunksFac128: PARODE_RUN2((kg_chunks_fac<128, false>), (kg_chunks_fac<128, true>), Chunks, 512); break;
    case kChunksTwo160: PARODE_RUN2((kg_chunks_two<160, false>), (kg_chunks_two<160, true>), Chunks, 320); break;
    case kChunksTwo192: PARODE_RUN2((kg_chunks_two<192, false>), (kg_chunks_two<192, true>), Chunks, 384); break;
    case kChunksTwo224: PARODE_RUN2((kg_chunks_two<224, false>), (kg_chunks_two<224, true>), Chunks, 448); break;
    case kChunksTwo256: PARODE_RUN2((kg_chunks_two<256, false>), (kg_chunks_two<256, true>), Chunks, 512); break;
    case kDense16: PARODE_RUN2((kg_chunks_dense<16, false>), (kg_chunks_dense<16, true>), Chunks, 64); break;
    case kDense32: PARODE_RUN2((kg_chunks_dense<32, false>), (kg_chunks_dense<32, true>), Chunks, 128); break;
    case kDense48: PARODE_RUN2((kg_chunks_dense<48, false>), (kg_chunks_dense<48, true>), Chunks, 192); break;
    case kDense64: PARODE_RUN2((kg_chunks_dense<64, false>), (kg_chunks_dense<64, true>), Chunks, 256); break;
    case kDense96: PARODE_RUN2((kg_chunks_dense<96, false>), (kg_chunks_dense<96, true>), Chunks, 384); break;
    case kDense128: PARODE_RUN2((kg_chunks_dense<128, false>), (kg_chunks_dense<128, true>), Chunks, 512); break;
    case kDense160: PARODE_RUN2((kg_chunks_dense<160, false>), (kg_chunks_dense<160, true>), Chunks, 640); break;
    case kDense192: PARODE_RUN2((kg_chunks_dense<192, false>), (kg_chunks_dense<192, true>), Chunks, 768); break;
    case kDense224: PARODE_RUN2((kg_chunks_dense<224, false>), (kg_chunks_dense<224, true>), Chunks, 896); break;
    case kDense256: PARODE_RUN2((kg_chunks_dense<256, false>), (kg_chunks_dense<256, true>), Chunks, 1024); break;
    case kLevel32: PARODE_RUN2((kg_level<float, false, false, true>), (kg_level<float, true, false, true>), Level, 256); break;
    case kLevel32W: PARODE_RUN2((kg_level<float, false, false, false>), (kg_level<float, true, false, false>), Level, 256); break;
    case kLevel32Delta: PARODE_RUN2((kg_level<float, false, true, true>), (kg_level<float, true, true, true>), Level, 256); break;
    case kLevel32DeltaW: PARODE_RUN2((kg_level<float, false, true, false>), (kg_level<float, true, true, false>), Level, 256); break;
    case kLevelG: PARODE_RUN2((kg_level_gemm<false, false>), (kg_level_gemm<true, false>), Level, 256); break;
    case kLevelGDelta: PARODE_RUN2((kg_level_gemm<false, true>), (kg_level_gemm<true, true>), Level, 256); break;
    case kLevel64: PARODE_RUN2((kg_level<double, false, false, false>), (kg_level<double, true, false, false>), Level, 256); break;
    case kGather: PARODE_RUN(kg_gather, Gather, 256); break;
    case kTopMfma: PARODE_RUN(kg_top<true>, Top, kTopThreads); break;
    case kTopSimt: PARODE_RUN(kg_top<false>, Top, kTopThreads); break;
    case kRoot: PARODE_RUN(kg_root, Root, 256); break;
    case kOutputsLanes: PARODE_RUN(kg_out_lanes, Outputs, kOutLanes); break;
    case kOutputsMfma: PARODE_RUN(kg_out_mfma, Outputs, 256); break;
    case kOutputsKc8: PARODE_RUN(kg_out_kc<8>, Outputs, kOutLanes); break;
    case kOutputsKc16: PARODE_RUN(kg_out_kc<16>, Outputs, kOutLanes); break;
    case kOutputsKc24: PARODE_RUN(kg_out_kc<24>, Outputs, kOutLanes); break;
    case kOutputsKc32: PARODE_RUN(kg_out_kc<32>, Outputs, kOutLanes); break;
    default: throw std::logic_error("grouped launch: unknown kernel");
    }
#undef PARODE_RUN2
#undef PARODE_RUN
    check_launch("grouped launch");
}

}  // namespace g

}  // namespace k
}  // namespace parode
