// Dispersion-spectrum kernels of the SWD hot path for gfx950 (MI355X):
// trace DFT at the Radon bins, the per-frequency complex slant stack and its
// adjoint (both on FP32 MFMA), spectrum normalisation, dispersion picking in
// wavefront code, and the fused SoftMax (soft-argmax) backward.
//
// Reference: core/RTr.m:5-32 (RTl.m, RTrAD.m, RTlAD.m), core/pickFDC1.m:19-131,
// core/LHDispPick.m:4-12, core/ADWDgrad_1.m:13-77.
//
// Column ("col") = one (shot, side) spectrum: col = side * nshots + shot,
// side 0 = right window (RTr), 1 = left window (RTl, traces read flipped).
// Frequency index k = 0..npair-1 is MATLAB column luoj = lf + k, i.e. DFT bin
// c = lf - 1 + k and kernel frequency df * c.
#include "swd_internal.h"
#include "swd_disp.h"

namespace swd {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16x16x4(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Twiddle table W[t][k2] (nt_eff x 2 npair): cos and -sin of 2 pi c t / ccn,
// c = lf - 1 + k, phase reduced exactly in integers (c t mod ccn).
// ---------------------------------------------------------------------------
__global__ void dft_table_kernel(float* W, int nt_eff, int npair, int c0, int ccn)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (k >= npair || t >= nt_eff) return;
    const long long r = ((long long)(c0 + k) * (long long)t) % ccn;
    double s, c;
    sincospi(2.0 * (double)r / (double)ccn, &s, &c);
    W[(size_t)t * 2 * npair + k] = (float)c;
    W[(size_t)t * 2 * npair + npair + k] = (float)(-s);
}

// ---------------------------------------------------------------------------
// Strided batched FP32 GEMM on MFMA 16x16x4:
//   C(b, m, n) = alpha * sum_k A(b, m, k) B(b, k, n)
// BM x BN tile per 256-thread block, 2 x 2 wavefronts of (BM/2) x (BN/2)
// (BM/32 x BN/32 MFMA tiles each); K tiles of BK double-buffered in LDS, the
// next tile's global loads in flight during the current tile's MFMAs, one
// barrier per tile.  Every shape runs the same MFMA chain per output element
// (ascending k, four at a time), so tile shapes and K tiles give bitwise
// equal results.  Measured on the production shapes (60 shots): the inverse
// DFT (M 6000, N 600, K 302) at 128 x 128, BK 16: 1.37 ms in the gradient,
// against 1.99 ms for the 64 x 64 single-buffer kernel it replaced; the DFT
// at the bins (M 302, N 600, K 6000) stays at 64 x 64, BK 32 (1.70 ms;
// 64 x 128 took 1.73 ms there, though a standalone run favoured it).
// ---------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 16;

// LDS layout of an operand tile, chosen at compile time by the operand's
// global layout so that the coalesced stores and the fragment reads (16 rows
// m at k .. k+3) are free of bank conflicts:
//   m (or n) contiguous in memory: [k][m], rows R + 16 floats apart (== 16 mod 64);
//   k contiguous in memory:        [m][k], rows BK + 2 floats apart.
template <int R, int BK, bool MF>
struct GemmLds {
    static constexpr int S = MF ? R + 16 : BK + 2;
    static constexpr int SIZE = MF ? BK * S : R * S;
    static __device__ __forceinline__ int off(int mm, int kk) { return MF ? kk * S + mm : mm * S + kk; }
};

template <int BM, int BN, int BK, bool AMF, bool BNF>
__global__ __launch_bounds__(256) void gemm_f32_db(GemmArgs g)
{
    using LA = GemmLds<BM, BK, AMF>;
    using LB = GemmLds<BN, BK, BNF>;
    constexpr int TM = BM / 32, TN = BN / 32, NA = BM * BK / 256, NB = BN * BK / 256;
    __shared__ float As[2][LA::SIZE];
    __shared__ float Bs[2][LB::SIZE];
    const int b = blockIdx.z;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const float* A = g.A + (size_t)b * g.sAb;
    const float* B = g.B + (size_t)b * g.sBb;
    const int wm = (w & 1) * (BM / 2), wn = (w >> 1) * (BN / 2);
    v4f acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    // element q of this thread's share of a tile: (row of m or n, k);
    // consecutive threads walk the operand's contiguous index
    auto a_at = [&](int q, int& mm, int& kk) {
        const int e = tid + 256 * q;
        if (AMF) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
    };
    auto b_at = [&](int q, int& nn, int& kk) {
        const int e = tid + 256 * q;
        if (BNF) { nn = e % BN; kk = e / BN; } else { kk = e % BK; nn = e / BK; }
    };
    float ra[NA], rb[NB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            int mm, kk;
            a_at(q, mm, kk);
            const int gm = m0 + mm, gk = k0 + kk;
            ra[q] = (gm < g.M && gk < g.K) ? A[(size_t)gm * g.sAm + (size_t)gk * g.sAk] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            int nn, kk;
            b_at(q, nn, kk);
            const int gn = n0 + nn, gk = k0 + kk;
            rb[q] = (gn < g.N && gk < g.K) ? B[(size_t)gk * g.sBk + (size_t)gn * g.sBn] : 0.f;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            int mm, kk;
            a_at(q, mm, kk);
            As[buf][LA::off(mm, kk)] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            int nn, kk;
            b_at(q, nn, kk);
            Bs[buf][LB::off(nn, kk)] = rb[q];
        }
    };
    gload(0);
    lstore(0);
    __syncthreads();
    int cur = 0;
    const int l16 = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + lk;
            float a[TM], bb[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[cur][LA::off(wm + i * 16 + l16, kr)];
#pragma unroll
            for (int j = 0; j < TN; ++j) bb[j] = Bs[cur][LB::off(wn + j * 16 + l16, kr)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = mfma16x16x4(a[i], bb[j], acc[i][j]);
        }
        if (more) lstore(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    float* C = g.C + (size_t)b * g.sCb;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + j * 16 + l16;
                if (gm < g.M && gn < g.N) C[(size_t)gm * g.sCm + (size_t)gn * g.sCn] = g.alpha * acc[i][j][r];
            }
}

template <int BM, int BN, int BK>
static void launch_gemm_db(hipStream_t st, const GemmArgs& g, int batch)
{
    const dim3 grid(div_up(g.N, BN), div_up(g.M, BM), batch);
    if (g.sAm == 1 && g.sBn == 1) gemm_f32_db<BM, BN, BK, true, true><<<grid, 256, 0, st>>>(g);
    else if (g.sAm == 1) gemm_f32_db<BM, BN, BK, true, false><<<grid, 256, 0, st>>>(g);
    else if (g.sBn == 1) gemm_f32_db<BM, BN, BK, false, true><<<grid, 256, 0, st>>>(g);
    else gemm_f32_db<BM, BN, BK, false, false><<<grid, 256, 0, st>>>(g);
}

// ---------------------------------------------------------------------------
// The same GEMM in BF16 split precision ("BF16 x 3") on the bf16 MFMA units:
// each FP32 operand is split into x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (round to nearest even), and a b = hi_a hi_b + hi_a lo_b
// + lo_a hi_b accumulates in FP32 (the lo_a lo_b term, ~2^-16 relative, is
// dropped).  The split happens once per element on the way into LDS, where
// both planes are stored k-contiguous so that one 8-byte read gives a lane
// its 4 k-values of v_mfma_f32_16x16x16_bf16 (A[m = l%16][k = 4(l/16)+i],
// B[k = 4(l/16)+i][n = l%16], D[m = 4(l/16)+r][n = l%16]).
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rn(float x)
{
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

constexpr int GKP = GBK + 4;  // LDS row of 16 k-values (+4 to spread banks)

__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(GemmArgs g)
{
    __shared__ unsigned short Ah[GBM][GKP], Al[GBM][GKP], Bh[GBN][GKP], Bl[GBN][GKP];
    const int b = blockIdx.z;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const float* A = g.A + (size_t)b * g.sAb;
    const float* B = g.B + (size_t)b * g.sBb;
    const bool a_mfast = g.sAm == 1;
    const bool b_nfast = g.sBn == 1;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    v4f acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4), lr = lane & 15;
    for (int k0 = 0; k0 < g.K; k0 += GBK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            int mm, kk;
            if (a_mfast) { mm = e % GBM; kk = e / GBM; } else { kk = e % GBK; mm = e / GBK; }
            const int gm = m0 + mm, gk = k0 + kk;
            const float av = (gm < g.M && gk < g.K) ? A[(size_t)gm * g.sAm + (size_t)gk * g.sAk] : 0.f;
            const unsigned short ah = bf16_rn(av);
            Ah[mm][kk] = ah;
            Al[mm][kk] = bf16_rn(av - bf16_to_f(ah));
            int nn;
            if (b_nfast) { nn = e % GBN; kk = e / GBN; } else { kk = e % GBK; nn = e / GBK; }
            const int gn = n0 + nn, gk2 = k0 + kk;
            const float bv = (gn < g.N && gk2 < g.K) ? B[(size_t)gk2 * g.sBk + (size_t)gn * g.sBn] : 0.f;
            const unsigned short bh = bf16_rn(bv);
            Bh[nn][kk] = bh;
            Bl[nn][kk] = bf16_rn(bv - bf16_to_f(bh));
        }
        __syncthreads();
        v4s ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ah[i] = *(const v4s*)&Ah[wm + i * 16 + lr][kq];
            al[i] = *(const v4s*)&Al[wm + i * 16 + lr][kq];
            bh[i] = *(const v4s*)&Bh[wn + i * 16 + lr][kq];
            bl[i] = *(const v4s*)&Bl[wn + i * 16 + lr][kq];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        __syncthreads();
    }
    float* C = g.C + (size_t)b * g.sCb;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < g.M && gn < g.N) C[(size_t)gm * g.sCm + (size_t)gn * g.sCn] = g.alpha * acc[i][j][r];
            }
}

// ---------------------------------------------------------------------------
// Per-trace statistics in one pass over the gathers traces[shot][t][g]:
//   peak[shot][g]     = max_t |x| over all nt samples (norm_trace's divisor,
//                       fieldexamples/outcrop/core/norm_trace.m:8-16; fmaxf
//                       skips NaN samples as MATLAB's max does);
//   par[shot][g][0/1] = sum of x over even / odd t (FP64; the fk-filter
//                       spectrum below needs them).
// Either output may be null.  64 traces per block (one per lane, coalesced
// rows of g); wave w of 4 takes t = w, w + 4, ..., so waves 0, 2 see even t
// and waves 1, 3 odd t; the four meet in LDS in a fixed order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void trace_stats_kernel(const float* __restrict__ traces, int nt, int ng,
                                                          float* __restrict__ peak, double* __restrict__ par)
{
    __shared__ float red[4][64];
    __shared__ double rs[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = blockIdx.x * 64 + lane;
    const int shot = blockIdx.y;
    float m0 = 0.f, m1 = 0.f;
    double s0 = 0.0, s1 = 0.0;
    if (g < ng) {
        const float* p = traces + (size_t)shot * nt * ng + g;
        int t = w;
        for (; t + 4 < nt; t += 8) {
            const float a = p[(size_t)t * ng], b = p[(size_t)(t + 4) * ng];
            m0 = fmaxf(m0, fabsf(a));
            m1 = fmaxf(m1, fabsf(b));
            s0 += (double)a;
            s1 += (double)b;
        }
        if (t < nt) {
            const float a = p[(size_t)t * ng];
            m0 = fmaxf(m0, fabsf(a));
            s0 += (double)a;
        }
    }
    red[w][lane] = fmaxf(m0, m1);
    rs[w][lane] = s0 + s1;
    __syncthreads();
    if (w == 0 && g < ng) {
        const size_t o = (size_t)shot * ng + g;
        if (peak) peak[o] = fmaxf(fmaxf(red[0][lane], red[1][lane]), fmaxf(red[2][lane], red[3][lane]));
        if (par) {
            par[2 * o] = rs[0][lane] + rs[2][lane];
            par[2 * o + 1] = rs[1][lane] + rs[3][lane];
        }
    }
}

// ---------------------------------------------------------------------------
// fk_filter (fieldexamples/QadimahFault/core/fk_filter.m:48-53,283-296, run by
// RTr/RTl when FK == 1 on each window X (nt x L) before the FFT) folded into
// the spectrum at the bins.  The filter keeps the (f < 0, k >= 0) and
// (f >= 0, k < 0) quadrants of the 2-D DFT of X zero-padded to N x M
// (N = 2 nt, M = 2 L), so it is real-linear and separable into one-sided
// filters along t and x.  With W the DFT at the bins (t < nt_eff), the
// spectrum of the filtered window is, column l of the window:
//   Df[:, l] = D[:, l]/2 - (1/M) sum_{l - l' odd} [ (2/N)(O Xe[l'] + E Xo[l'])
//                                                 + 2 cot(pi (l - l')/M) H[:, l'] ]
// where D = W^T X, E / O = sums of W over even / odd t, Xe / Xo = sums of each
// trace over even / odd t, and H = U' X with
//   U'[k2][s] = -(1/N) sum_{t < nt_eff, t - s odd} W[t][k2] cot(pi (t - s)/N)
// (the discrete Hilbert kernel of the one-sided time filter).  D and H are
// two GEMMs over the whole gathers; the window kernel below mixes each
// window's traces.
// ---------------------------------------------------------------------------

// cotab[tau] = cot(pi tau / N) for tau = 1 .. n-1 (FP64); cotab[0] = 0
__global__ void cot_table_kernel(double* cotab, int n, double N)
{
    const int tau = blockIdx.x * blockDim.x + threadIdx.x;
    if (tau >= n) return;
    if (tau == 0) { cotab[0] = 0.0; return; }
    double sv, cv;
    sincospi((double)tau / N, &sv, &cv);
    cotab[tau] = cv / sv;
}

// U'[s][k2] (layout of W: [s][2 npair]) in FP64 from W[t][k2], s < nt.
__global__ __launch_bounds__(64) void fk_hilbert_table_kernel(const float* __restrict__ W, int nt_eff, int nt,
                                                              int npair2, const double* __restrict__ cotab,
                                                              float* __restrict__ Up)
{
    const int k2 = blockIdx.x * 64 + threadIdx.x;
    const int s = blockIdx.y;
    if (k2 >= npair2) return;
    double acc = 0.0;
    for (int t = (s + 1) & 1; t < nt_eff; t += 2) {
        const int tau = t - s;
        const double c = tau > 0 ? cotab[tau] : -cotab[-tau];
        acc += (double)W[(size_t)t * npair2 + k2] * c;
    }
    Up[(size_t)s * npair2 + k2] = (float)(-acc / (2.0 * nt));
}

// EO[0][k2] = sum_{t even < nt_eff} W[t][k2], EO[1][k2] = the odd-t sum (FP64)
__global__ void fk_parity_table_kernel(const float* __restrict__ W, int nt_eff, int npair2, double* __restrict__ EO)
{
    const int k2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (k2 >= npair2) return;
    double e = 0.0, o = 0.0;
    for (int t = 0; t < nt_eff; t += 2) e += (double)W[(size_t)t * npair2 + k2];
    for (int t = 1; t < nt_eff; t += 2) o += (double)W[(size_t)t * npair2 + k2];
    EO[k2] = e;
    EO[npair2 + k2] = o;
}

constexpr int FKK = 8;      // frequencies per block (16 rows of the real representation)
constexpr int FKL = 256;    // window positions per block and per staged chunk
constexpr int kFkMaxL = 2048;

// X[k][l][col] = filtered spectrum of window col at frequency k, position l
// (0 beyond L, as the plain gather).  With trace peaks, every trace enters
// divided by its peak (norm_trace before the window).  Block: one col, FKK
// frequencies, FKL positions; the col's H rows are staged through LDS in
// chunks of FKL positions, the cot weights of the window in LDS.
__global__ __launch_bounds__(FKL) void fk_window_kernel(const float* __restrict__ D, const float* __restrict__ H,
                                                        const double* __restrict__ par, const float* __restrict__ peak,
                                                        const double* __restrict__ EO, const DispCol* __restrict__ cols,
                                                        int ncol, int npair, int ng, int Lmax, int nt, float2* X)
{
    __shared__ float Hs[2 * FKK][FKL];
    __shared__ float ctab[kFkMaxL];
    __shared__ double red[4][FKL / 64];
    const int col = blockIdx.x;
    const int k0 = blockIdx.y * FKK;
    const int tid = threadIdx.x;
    const int l = blockIdx.z * FKL + tid;
    const DispCol c = cols[col];
    const int npair2 = 2 * npair;
    const int L = c.enabled ? c.L : 0;
    if (L == 0 || blockIdx.z * FKL >= L) {
        if (l < Lmax)
            for (int kk = 0; kk < FKK && k0 + kk < npair; ++kk)
                X[((size_t)(k0 + kk) * Lmax + l) * ncol + col] = make_float2(0.f, 0.f);
        return;  // uniform over the block
    }
    const double M = 2.0 * L, N = 2.0 * nt;
    const float* Dc = D + (size_t)c.shot * npair2 * ng;
    const float* Hc = H + (size_t)c.shot * npair2 * ng;
    auto qof = [&](int lp) -> float { return peak ? 1.0f / peak[(size_t)c.shot * ng + c.tr0 + c.dir * lp] : 1.0f; };
    // cot(pi tau / M), tau = 0 .. L-1, and the parity-class sums of the (scaled) trace sums
    for (int tau = tid; tau < L; tau += FKL) {
        double sv, cv;
        sincospi((double)tau / M, &sv, &cv);
        ctab[tau] = tau ? (float)(cv / sv) : 0.f;
    }
    double se0 = 0.0, se1 = 0.0, so0 = 0.0, so1 = 0.0;
    for (int lp = tid; lp < L; lp += FKL) {
        const size_t o = 2 * ((size_t)c.shot * ng + c.tr0 + c.dir * lp);
        const double q = (double)qof(lp);
        const double xe = par[o] * q, xo = par[o + 1] * q;
        if (lp & 1) { se1 += xe; so1 += xo; } else { se0 += xe; so0 += xo; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        se0 += __shfl_xor(se0, off); se1 += __shfl_xor(se1, off);
        so0 += __shfl_xor(so0, off); so1 += __shfl_xor(so1, off);
    }
    if ((tid & 63) == 0) { red[0][tid >> 6] = se0; red[1][tid >> 6] = se1; red[2][tid >> 6] = so0; red[3][tid >> 6] = so1; }
    float acc[2 * FKK];
#pragma unroll
    for (int r = 0; r < 2 * FKK; ++r) acc[r] = 0.f;
    const int par_l = l & 1;
    for (int c0 = 0; c0 < L; c0 += FKL) {
        __syncthreads();
        for (int e = tid; e < 2 * FKK * FKL; e += FKL) {
            const int r = e / FKL, j = e % FKL;
            const int k2 = r < FKK ? k0 + r : npair + k0 + (r - FKK);
            const int lp = c0 + j;
            float v = 0.f;
            if (lp < L && k0 + (r % FKK) < npair) v = Hc[(size_t)k2 * ng + c.tr0 + c.dir * lp] * qof(lp);
            Hs[r][j] = v;
        }
        __syncthreads();
        if (l < L) {
            const int jend = min(FKL, L - c0);
            for (int j = (par_l + 1 + c0) & 1; j < jend; j += 2) {
                const int tau = l - (c0 + j);
                const float cw = tau > 0 ? ctab[tau] : -ctab[-tau];
#pragma unroll
                for (int r = 0; r < 2 * FKK; ++r) acc[r] = fmaf(cw, Hs[r][j], acc[r]);
            }
        }
    }
    if (l >= Lmax) return;
    if (l >= L) {
        for (int kk = 0; kk < FKK && k0 + kk < npair; ++kk)
            X[((size_t)(k0 + kk) * Lmax + l) * ncol + col] = make_float2(0.f, 0.f);
        return;
    }
    // sums over the parity class opposite to l
    const double se = par_l ? red[0][0] + red[0][1] + red[0][2] + red[0][3] : red[1][0] + red[1][1] + red[1][2] + red[1][3];
    const double so = par_l ? red[2][0] + red[2][1] + red[2][2] + red[2][3] : red[3][0] + red[3][1] + red[3][2] + red[3][3];
    const int g = c.tr0 + c.dir * l;
    const float ql = qof(l);
    const double cc = 2.0 / (M * N);
    for (int kk = 0; kk < FKK && k0 + kk < npair; ++kk) {
        const int kr = k0 + kk, ki = npair + k0 + kk;
        const double re = 0.5 * (double)(Dc[(size_t)kr * ng + g] * ql) - cc * (EO[npair2 + kr] * se + EO[kr] * so) -
                          (2.0 / M) * (double)acc[kk];
        const double im = 0.5 * (double)(Dc[(size_t)ki * ng + g] * ql) - cc * (EO[npair2 + ki] * se + EO[ki] * so) -
                          (2.0 / M) * (double)acc[FKK + kk];
        X[((size_t)kr * Lmax + l) * ncol + col] = make_float2((float)re, (float)im);
    }
}

// ---------------------------------------------------------------------------
// Peak normalisation of each shot's adjoint source (the field scripts'
// ADWDgrad_1.m:79, a_data_res./max(abs(a_data_res(:)))): kPeakParts blocks
// per shot reduce strided shares to partial maxima, then every block of the
// scaling pass combines the partials of its shot and divides its share.  An
// all-zero source gives 0/0 = NaN, as the division in the script.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void shot_peak_part_kernel(const float* __restrict__ a, size_t n,
                                                             float* __restrict__ part)
{
    __shared__ float red[4];
    const float* p = a + (size_t)blockIdx.y * n;
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kPeakParts * 256)
        m = fmaxf(m, fabsf(p[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        part[(size_t)blockIdx.y * kPeakParts + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ __launch_bounds__(256) void shot_peak_scale_kernel(float* __restrict__ a, size_t n,
                                                              const float* __restrict__ part)
{
    __shared__ float pk;
    if (threadIdx.x < 64) {
        float m = part[(size_t)blockIdx.y * kPeakParts + threadIdx.x];
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (threadIdx.x == 0) pk = m;
    }
    __syncthreads();
    const float q = pk;
    float* p = a + (size_t)blockIdx.y * n;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kPeakParts * 256) p[i] = p[i] / q;
}

// ---------------------------------------------------------------------------
// Window gather: X[k][l][col] = D(shot, k, trace(col, l)), 0 beyond L(col).
// D layout: D[shot][k2][g] with Re at k2 = k and Im at k2 = npair + k.
// With trace peaks (norm_trace), each spectrum is divided by its trace's
// peak: the DFT is linear, so this is the spectrum of the normalised trace.
// A zero trace gives 0/0 = NaN, as norm_trace does.
// ---------------------------------------------------------------------------
__global__ void radon_gather_kernel(const float* D, const DispCol* cols, int ncol, int npair, int ng, int Lmax,
                                    const float* peak, float2* X)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int k = blockIdx.z;
    if (col >= ncol) return;
    const DispCol c = cols[col];
    float2 v = make_float2(0.f, 0.f);
    if (c.enabled && l < c.L) {
        const int g = c.tr0 + c.dir * l;
        const float* d = D + (size_t)c.shot * 2 * npair * ng;
        v.x = d[(size_t)k * ng + g];
        v.y = d[(size_t)(npair + k) * ng + g];
        if (peak) {
            const float q = peak[(size_t)c.shot * ng + g];
            v.x = v.x / q;
            v.y = v.y / q;
        }
    }
    X[((size_t)k * Lmax + l) * ncol + col] = v;
}

// ---------------------------------------------------------------------------
// Complex slant stack with the Radon kernel generated in registers.
// Forward (ADJ = false), for each frequency k:
//   OUT[k][j][col] = sum_l exp(+i phi(j,l)) IN[k][l][col]     (j < np, l < Lmax)
// Adjoint (ADJ = true):
//   OUT[k][l][col] = sum_j exp(-i phi(j,l)) IN[k][j][col]     (l < Lmax, j < np)
// phi(j,l) = 2 pi * frac(f_k * p_j * x_l), f_k = df (c0 + k), p_j = 1/(vmin + j),
// x_l = (m + l) ds.  The phase is reduced in double, sin/cos taken in float.
// 64 (rows) x 64 (cols) tile per block, 4 real MFMA 16x16x4 per complex step.
// ---------------------------------------------------------------------------
template <bool ADJ>
__global__ __launch_bounds__(256) void slant_stack_kernel(SlantArgs a)
{
    __shared__ float Ar[GBK][GBM + 4], Ai[GBK][GBM + 4];
    __shared__ float Br[GBK][GBN + 4], Bi[GBK][GBN + 4];
    const int k = blockIdx.z;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int M = ADJ ? a.Lmax : a.np;      // output rows
    const int K = ADJ ? a.np : a.Lmax;      // contraction length
    const double fk = a.df * (double)(a.c0 + k);
    const float2* IN = a.in + (size_t)k * K * a.ncol;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    v4f cr[2][2], ci[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) { cr[i][j] = v4f{0.f, 0.f, 0.f, 0.f}; ci[i][j] = cr[i][j]; }
    for (int k0 = 0; k0 < K; k0 += GBK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            // generated operand: row mm (output row), contraction kk
            const int mm = e % GBM, kk = e / GBM;
            const int gm = m0 + mm, gk = k0 + kk;
            float sr = 0.f, si = 0.f;
            if (gm < M && gk < K) {
                const int j = ADJ ? gk : gm;
                const int l = ADJ ? gm : gk;
                const double p = 1.0 / (a.vmin + (double)j);
                const double x = (double)(a.m + l) * a.ds;
                double u = fk * (p * x);
                u -= floor(u);
                float s, c;
                sincospif(2.0f * (float)u, &s, &c);
                sr = c;
                si = ADJ ? -s : s;
            }
            Ar[kk][mm] = sr;
            Ai[kk][mm] = si;
            // memory operand: contraction kk2, column nn (col fastest in memory)
            const int nn = e % GBN, kk2 = e / GBN;
            const int gn = n0 + nn, gk2 = k0 + kk2;
            float2 v = make_float2(0.f, 0.f);
            if (gn < a.ncol && gk2 < K) v = IN[(size_t)gk2 * a.ncol + gn];
            Br[kk2][nn] = v.x;
            Bi[kk2][nn] = v.y;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            const int kr = kk + (lane >> 4);
            float ar[2], ai[2], br[2], bi[2];
            for (int i = 0; i < 2; ++i) {
                ar[i] = Ar[kr][wm + i * 16 + (lane & 15)];
                ai[i] = Ai[kr][wm + i * 16 + (lane & 15)];
                br[i] = Br[kr][wn + i * 16 + (lane & 15)];
                bi[i] = Bi[kr][wn + i * 16 + (lane & 15)];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    cr[i][j] = mfma16x16x4(ar[i], br[j], cr[i][j]);
                    cr[i][j] = mfma16x16x4(-ai[i], bi[j], cr[i][j]);
                    ci[i][j] = mfma16x16x4(ar[i], bi[j], ci[i][j]);
                    ci[i][j] = mfma16x16x4(ai[i], br[j], ci[i][j]);
                }
        }
        __syncthreads();
    }
    float2* OUT = a.out + (size_t)k * M * a.ncol;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < a.ncol) OUT[(size_t)gm * a.ncol + gn] = make_float2(cr[i][j][r], ci[i][j][r]);
            }
}

// ---------------------------------------------------------------------------
// |mm|, column maximum, normalised spectrum (RTr.m:28-32, RTrAD.m:27-32).
// One block per (col, k); outputs ml[col][k][j], mlr[col][k][j].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void radon_abs_kernel(const float2* MM, int np, int npair, int ncol, float* ml,
                                                        float* mlr)
{
    const int col = blockIdx.x, k = blockIdx.y;
    __shared__ float red[256];
    float mx = 0.f;
    bool nan = false;
    for (int j = threadIdx.x; j < np; j += 256) {
        const float2 v = MM[((size_t)k * np + j) * ncol + col];
        const float a = sqrtf(v.x * v.x + v.y * v.y);
        mlr[((size_t)col * npair + k) * np + j] = a;
        if (a != a) nan = true;
        else mx = fmaxf(mx, a);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const float cmax = red[0];
    (void)nan;
    for (int j = threadIdx.x; j < np; j += 256) {
        const size_t o = ((size_t)col * npair + k) * np + j;
        ml[o] = mlr[o] / cmax;  // 0/0 -> NaN for an all-zero column, as in MATLAB
    }
}

// ---------------------------------------------------------------------------
// pickFDC1 (pickFDC1.m:19-131) + LHDispPick (LHDispPick.m:4-12), one
// wavefront per column.  findpeaks(-y) local minima: a sample starting a run
// of equal values whose previous sample is larger and whose run is followed
// by a larger sample (MATLAB findpeaks defaults; endpoints never count).
// ---------------------------------------------------------------------------
// Wave-wide argmax by DPP: row_shr 1, 2, 4, 8 leave each row's (value,
// index) maximum in its lane 15, row_bcast 15 and 31 carry rows 0-2 into row
// 3, so lane 63 holds the wave's; a lane whose DPP source is invalid keeps its
// own pair.  Ties go to the smaller index, NaN never enters.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_keep(int x) { return __builtin_amdgcn_update_dpp(x, x, CTRL, RM, 0xf, false); }
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_keep(float x) { return __int_as_float(dpp_keep<CTRL, RM>(__float_as_int(x))); }
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_keep(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_keep<CTRL, RM>((int)(b & 0xffffffffLL)), hi = dpp_keep<CTRL, RM>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int RM, typename V>
__device__ __forceinline__ void argmax_dpp_step(V& best, int& bi)
{
    const V ob = dpp_keep<CTRL, RM>(best);
    const int oi = dpp_keep<CTRL, RM>(bi);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
}
// the index of the wave's maximum (uniform)
template <typename V>
__device__ __forceinline__ int wave_argmax_index(V best, int bi)
{
    argmax_dpp_step<0x111, 0xf>(best, bi);  // row_shr:1
    argmax_dpp_step<0x112, 0xf>(best, bi);  // row_shr:2
    argmax_dpp_step<0x114, 0xf>(best, bi);  // row_shr:4
    argmax_dpp_step<0x118, 0xf>(best, bi);  // row_shr:8
    argmax_dpp_step<0x142, 0xa>(best, bi);  // row_bcast:15 into rows 1 and 3
    argmax_dpp_step<0x143, 0xc>(best, bi);  // row_bcast:31 into rows 2 and 3
    return __builtin_amdgcn_readlane(bi, 63);
}

// is position p (0-based within v[lo..hi]) a findpeaks location of -v ?
template <typename V>
__device__ bool is_local_min(const V* v, int lo, int hi, int p)
{
    const int i = lo + p;
    const V y = v[i];
    if (y != y) return false;
    if (i == lo) return false;
    const V yp = v[i - 1];
    if (!(yp > y)) return false;  // previous (last of previous run) must be larger (NaN fails)
    int e = i;
    while (e + 1 <= hi && v[e + 1] == y) ++e;
    if (e + 1 > hi) return false;
    const V yn = v[e + 1];
    return yn > y;
}

// lp = 0 picks from the column of the spectrum's largest |mm| (the field
// scripts' LHDispPick: [~,ini] = max(max(mlr))).  The misfit averages the
// columns r0..npair-1-r1 (0-based; the field scripts' cr_pre(a:end-b)).
// V is the spectrum's type: float for the engine and swd_lhdisppick, double
// for callers that pick on FP64 spectra (swd_lhdisppick_f64).
//
// The tracking is sequential over the frequency columns, so one wavefront
// per (shot, side) spends most of a step waiting for its next column.  The
// columns ahead stream through a ring of kPickRing register buffers
// (NK values per lane each), issued kPickRing - 1 steps before the column is
// staged in LDS; every load is unconditional (indices clamped into the
// spectrum), so the compiler's wait counts are exact and no wait reaches
// past the column that is needed.
constexpr int kPickRing = 4;

template <typename V, int NK>
__global__ __launch_bounds__(64) void pick_kernel(const V* ml, const V* mlr, const DispCol* cols, int np,
                                                  int npair, int lp, float fg, float lfg, double vmin, int r0, int r1,
                                                  const double* cr0, double* cr_pre, double* res, int* status)
{
    const int col = blockIdx.x;
    const int lane = threadIdx.x;
    const DispCol c = cols[col];
    __shared__ int idx[kMaxPair], dvmax[kMaxPair], dvmin[kMaxPair], chz[kMaxPair];
    __shared__ float qj[kMaxPair];
    __shared__ double fidx[kMaxPair];
    // the frequency column being tracked and the next one, staged in LDS
    __shared__ V colv[2][kPickMaxNp];
    if (!c.enabled) return;
    const V* base = ml + (size_t)col * npair * np;
    const int a = np;
    V pf[kPickRing][NK];
    // column ci (clamped into 0..npair-1) into ring buffer u; the element
    // index is clamped too, so every lane issues NK loads
    auto fetch = [&](V* buf, int ci) {
        const V* cb = base + (size_t)min(max(ci, 0), npair - 1) * np;
#pragma unroll
        for (int k = 0; k < NK; ++k) buf[k] = cb[min(k * 64 + lane, np - 1)];
    };
    auto put = [&](const V* buf, int b) {
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int j = k * 64 + lane;
            if (j < np) colv[b][j] = buf[k];
        }
    };

    // argmax over v[lo..hi] (NaN ignored, first index of the maximum); 1-based result
    auto argmax = [&](const V* v, int lo, int hi) -> int {
        V best = -INFINITY;
        int bi = 0x7fffffff;
        for (int i = lo + lane; i <= hi; i += 64) {
            const V x = v[i];
            if (x == x && (x > best || (x == best && i < bi))) { best = x; bi = i; }
        }
        bi = wave_argmax_index(best, bi);
        return bi == 0x7fffffff ? 1 : bi - lo + 1;
    };
    if (lp == 0) {
        // column maxima of |mm| (NaN ignored), then the first column holding the largest
        const V* u = mlr + (size_t)col * npair * np;
        V best = -INFINITY;
        int bk = 0;
        for (int k = 0; k < npair; ++k) {
            V m = -INFINITY;
            for (int j = lane; j < np; j += 64) m = fmax(m, u[(size_t)k * np + j]);
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
            if (m > best) { best = m; bk = k; }
        }
        lp = bk + 1;
    }
    const int L1 = lp - 1;  // 0-based index of lp
    if (lane == 0) for (int i = 0; i < npair; ++i) { chz[i] = 1; qj[i] = NAN; dvmax[i] = 0; dvmin[i] = 0; }
    fetch(pf[0], L1);
    put(pf[0], 0);
    __syncthreads();
    const int cur = argmax(colv[0], 0, a - 1);
    if (lane == 0) idx[L1] = cur;
    __syncthreads();

    // lane k's 64-bit word (k uniform), through scalar reads
    auto lane_word = [](unsigned long long m, int k) -> unsigned long long {
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)(m & 0xffffffffull), k);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(m >> 32), k);
        return ((unsigned long long)hi << 32) | lo;
    };
    // The local minima of the whole column (findpeaks of -v over v[0..a-1]):
    // lane k < NK holds the bits of positions 64k .. 64k+63.  The windows of
    // a step are read off this mask:
    //  - first_min(v, ip-1, a-1) finds the first position i >= ip that is a
    //    minimum of the column (position ip-1 itself is excluded, and the
    //    window ends where the column does, so nothing else differs);
    //  - last_min(v, 0, ip-1) finds the last minimum i <= ip-1 whose run of
    //    equal values ends before ip-1.  Only the last minimum of the column
    //    at or below ip-1 can have a run crossing ip-1 (an earlier one ends
    //    below the next minimum), so at most one candidate is skipped.
    // Built without branches: each lane reads the left, centre and right
    // values of its NK positions (independent LDS reads, clamped indices) and
    // tests for a strict minimum, which is the findpeaks rule wherever the
    // right neighbour differs.  A position whose right neighbour is equal
    // starts a run of equal values; those (rare in spectra) are decided by
    // the full rule.
    auto min_mask = [&](const V* v) -> unsigned long long {
        unsigned long long mine = 0, runs = 0;
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int p = k * 64 + lane;
            const int pc = min(p, a - 1);
            const V y = v[pc], yl = v[max(pc - 1, 0)], yr = v[min(pc + 1, a - 1)];
            const bool in = p >= 1 && p <= a - 2;  // the first and last positions are never minima
            const unsigned long long m = __ballot(in && yl > y && yr > y);
            const unsigned long long r = __ballot(in && yl > y && yr == y);
            if (lane == k) { mine = m; runs = r; }
        }
        if (__ballot(runs != 0ull)) {
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const unsigned long long rk = lane_word(runs, k);
                if (rk) {
                    const int p = k * 64 + lane;
                    const unsigned long long m = __ballot(((rk >> lane) & 1ull) && is_local_min(v, 0, a - 1, p));
                    if (lane == k) mine |= m;
                }
            }
        }
        return mine;
    };
    // first set position >= s, -1 if none
    auto first_from = [&](unsigned long long mine, int s) -> int {
        const int w = s >> 6;
        unsigned long long m = (lane < w || lane >= NK) ? 0ull : mine;
        if (lane == w) m &= ~0ull << (s & 63);
        const unsigned long long nz = __ballot(m != 0ull);
        if (!nz) return -1;
        const int k = __ffsll((long long)nz) - 1;
        const unsigned long long mk = lane_word(m, k);
        return k * 64 + __ffsll((long long)mk) - 1;
    };
    // last set position <= t, -1 if none
    auto last_upto = [&](unsigned long long mine, int t) -> int {
        if (t < 0) return -1;
        const int w = t >> 6;
        unsigned long long m = (lane > w || lane >= NK) ? 0ull : mine;
        if (lane == w && (t & 63) != 63) m &= (1ull << ((t & 63) + 1)) - 1ull;
        const unsigned long long nz = __ballot(m != 0ull);
        if (!nz) return -1;
        const int k = 63 - __clzll((long long)nz);
        const unsigned long long mk = lane_word(m, k);
        return k * 64 + 63 - __clzll((long long)mk);
    };

    // The tracking state of the last column (its idx, qj, dvmax, dvmin),
    // carried in registers from step to step; lane 0 also records each
    // column's values in LDS for the interpolation at the end.
    struct Track {
        int ip;
        float q;
        int dmax, dmin;
    };
    // one tracking step at column i from its neighbour's state t
    // (left sweep: dir = -1, width jumps above lfg rejected; right: dir = +1, fg)
    auto step = [&](const V* v, int i, int dir, const Track t) __attribute__((always_inline)) -> Track {
        const int ip = t.ip;
        if (dir < 0 && ip == a) {
            // carried over; qj[i] stays NaN and chz[i] stays 1
            if (lane == 0) { idx[i] = a; dvmax[i] = t.dmax; dvmin[i] = t.dmin; }
            return Track{a, NAN, t.dmax, t.dmin};
        }
        const unsigned long long mine = (ip > a - 3 && ip < 3) ? 0ull : min_mask(v);
        int dmax, dmin;
        if (ip > a - 3) dmax = a;
        else {
            // first_min(v, ip - 1, a - 1): 1-based location in the window
            const int m = first_from(mine, ip);
            dmax = m >= 0 ? ip + (m - (ip - 1) + 1) : a;
        }
        if (ip < 3) dmin = 1;
        else {
            // last_min(v, 0, ip - 1)
            const int hi = ip - 1;
            int m = last_upto(mine, hi);
            if (m >= 0) {
                const V y = v[m];
                int e = m;
                while (e + 1 <= hi && v[e + 1] == y) ++e;
                if (e + 1 > hi) m = last_upto(mine, m - 1);
            }
            dmin = m >= 0 ? m + 1 : 1;
        }
        const float q = (float)(dmax - dmin);
        const float qprev = t.q;
        if (dir < 0 ? q > qprev * lfg : q > fg * qprev) {
            const float qc = (float)(t.dmax - t.dmin);
            if (lane == 0) { chz[i] = 0; dvmax[i] = t.dmax; dvmin[i] = t.dmin; qj[i] = qc; idx[i] = ip; }
            return Track{ip, qc, t.dmax, t.dmin};
        }
        const int newidx = dmin + argmax(v, dmin - 1, dmax - 1) - 1;
        if (lane == 0) { idx[i] = newidx; dvmax[i] = dmax; dvmin[i] = dmin; qj[i] = q; }
        return Track{newidx, q, dmax, dmin};
    };
    // columns c0, c0 + dir, ... (n of them): column j is fetched into ring
    // buffer j % kPickRing at step j - kPickRing (or before the sweep) and
    // staged in LDS at the end of step j - 1
    auto sweep = [&](int c0, int dir, int n) __attribute__((always_inline)) {
        if (n <= 0) return;
        // the start column: idx = the argmax, dvmax = dvmin = 0, qj = NaN
        Track t{cur, NAN, 0, 0};
#pragma unroll
        for (int u = 0; u < kPickRing; ++u) fetch(pf[u], c0 + dir * min(u, n - 1));
        put(pf[0], 0);
        __syncthreads();
        // Every sub-step stages and fetches a column, also past the end of the
        // sweep (clamped, unused), so that each issues the same loads: the
        // compiler then waits only for the column being staged, which was
        // requested kPickRing - 1 steps earlier.  A tracking step guarded by
        // the same condition as its loads would let it assume the shortest
        // path and wait for the loads just issued.
        int b = 0;
        for (int j0 = 0; j0 < n; j0 += kPickRing) {
#pragma unroll
            for (int u = 0; u < kPickRing; ++u) {
                const int j = j0 + u;
                if (j < n) t = step(colv[b], c0 + dir * j, dir, t);
                __syncthreads();
                // column j + 1 into LDS; column j + kPickRing into the freed buffer
                put(pf[(u + 1) % kPickRing], b ^ 1);
                fetch(pf[u], c0 + dir * min(j + kPickRing, n - 1));
                __syncthreads();
                b ^= 1;
            }
        }
    };
    sweep(L1 - 1, -1, L1);               // left of lp
    sweep(L1 + 1, 1, npair - 1 - L1);   // right of lp
    // interp1(que, idx(que), 1:b, 'linear') then nearest extrapolation
    if (lane == 0) {
        int nq = 0, qfirst = -1, qlast = -1;
        for (int i = 0; i < npair; ++i)
            if (chz[i]) { ++nq; if (qfirst < 0) qfirst = i; qlast = i; }
        int st = 0;
        if (nq < 2) st = kErrPick;
        int prev = -1;
        for (int i = 0; i < npair; ++i) {
            if (chz[i]) { fidx[i] = (double)idx[i]; prev = i; continue; }
            if (i < qfirst || i > qlast) continue;
            int nxt = i + 1;
            while (!chz[nxt]) ++nxt;
            const double t = (double)(i - prev) / (double)(nxt - prev);
            fidx[i] = (double)idx[prev] + t * (double)(idx[nxt] - idx[prev]);
        }
        for (int i = 0; i < npair; ++i) {
            if (nq == 0) { fidx[i] = (double)idx[i]; continue; }
            if (i < qfirst) fidx[i] = (double)idx[qfirst];
            if (i > qlast) fidx[i] = (double)idx[qlast];
        }
        double acc = 0.0;
        for (int i = 0; i < npair; ++i) {
            const double v = fidx[i] + vmin;
            cr_pre[(size_t)col * npair + i] = v;
            const double d = v - cr0[(size_t)col * npair + i];
            if (i >= r0 && i < npair - r1) acc += d * d;
        }
        res[col] = acc / (double)max(npair - r1 - r0, 0);  // mean of an empty window is 0/0 = NaN
        status[col] = st;
    }
}

// LHDispPick method 2 (LHDispPick.m:8-9): column argmax (first index, NaN ignored) + vmin.
template <typename V>
__global__ __launch_bounds__(64) void argmax_pick_kernel(const V* ml, int np, int npair, double vmin,
                                                         const double* cr0, double* cr_pre)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const V* v = ml + (size_t)k * np;
    V best = -INFINITY;
    int bi = 0x7fffffff;
    for (int i = lane; i < np; i += 64) {
        const V x = v[i];
        if (x == x && (x > best || (x == best && i < bi))) { best = x; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const V ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (lane == 0) cr_pre[k] = (double)(bi == 0x7fffffff ? 1 : bi + 1) + vmin;
}

// ---------------------------------------------------------------------------
// Soft-argmax backward + phase (ADWDgrad_1.m:21-34), one block per (col, k):
//   G = mlr^(1/40);  E = softmax(beta G) (log-sum-exp in double);
//   kappa = sum_j j E_j;  g = cr_pre - cr0 (0 -> 1);
//   gml_j = 2 beta g E_j (j - kappa);  GMM[k][j][col] = (mm/|mm|) gml_j (0 when |mm| = 0)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void softargmax_kernel(const float* mlr, const float2* MM, const DispCol* cols,
                                                         int np, int npair, int ncol, double beta,
                                                         const double* cr_pre, const double* cr0, float2* GMM)
{
    const int col = blockIdx.x, k = blockIdx.y;
    const DispCol c = cols[col];
    __shared__ double red[256];
    __shared__ double red2[256];
    const float* G = mlr + ((size_t)col * npair + k) * np;
    const size_t ostride = (size_t)ncol;
    float2* out = GMM + (size_t)k * np * ncol + col;
    if (!c.enabled) {
        for (int j = threadIdx.x; j < np; j += 256) out[j * ostride] = make_float2(0.f, 0.f);
        return;
    }
    double mx = -INFINITY;
    for (int j = threadIdx.x; j < np; j += 256) mx = fmax(mx, beta * pow((double)G[j], 1.0 / 40.0));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double gmax = red[0];
    __syncthreads();
    double se = 0.0, sj = 0.0;
    for (int j = threadIdx.x; j < np; j += 256) {
        const double e = exp(beta * pow((double)G[j], 1.0 / 40.0) - gmax);
        se += e;
        sj += (double)(j + 1) * e;
    }
    red[threadIdx.x] = se;
    red2[threadIdx.x] = sj;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) { red[threadIdx.x] += red[threadIdx.x + s]; red2[threadIdx.x] += red2[threadIdx.x + s]; }
        __syncthreads();
    }
    const double Z = red[0];
    const double kappa = red2[0] / Z;
    double g = cr_pre[(size_t)col * npair + k] - cr0[(size_t)col * npair + k];
    if (g == 0.0) g = 1.0;
    for (int j = threadIdx.x; j < np; j += 256) {
        const double E = exp(beta * pow((double)G[j], 1.0 / 40.0) - gmax) / Z;
        const double gml = ((E * (double)(j + 1) - E * kappa) * 2.0 * g) * beta;
        const float2 m = MM[((size_t)k * np + j) * ncol + col];
        const double am = sqrt((double)m.x * m.x + (double)m.y * m.y);
        float2 o = make_float2(0.f, 0.f);
        if (am > 0.0) o = make_float2((float)((double)m.x / am * gml), (float)((double)m.y / am * gml));
        out[j * ostride] = o;
    }
}

// ---------------------------------------------------------------------------
// Placement of the adjoint slant stack into the inverse-DFT operand:
//   Bm[shot][g][k] = Re GD[k][l][col], Bm[shot][g][npair + k] = Im GD[k][l][col]
// with g = trace(col, l); the left side skips l = 0 (ADWDgrad_1.m:76).
// ---------------------------------------------------------------------------
__global__ void adj_place_kernel(const float2* GD, const DispCol* cols, int ncol, int npair, int ng, int Lmax,
                                 float* Bm)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y, col = blockIdx.z;
    if (k >= npair) return;
    const DispCol c = cols[col];
    if (!c.enabled || l >= c.L || (c.dir < 0 && l == 0)) return;
    const int g = c.tr0 + c.dir * l;
    const float2 v = GD[((size_t)k * Lmax + l) * ncol + col];
    float* o = Bm + ((size_t)c.shot * ng + g) * 2 * npair;
    o[k] = v.x;
    o[npair + k] = v.y;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int launch_dft_table(hipStream_t st, float* W, int nt_eff, int npair, int c0, int ccn)
{
    dft_table_kernel<<<dim3(div_up(npair, 128), nt_eff), 128, 0, st>>>(W, nt_eff, npair, c0, ccn);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_gemm_f32(hipStream_t st, const GemmArgs& g, int batch)
{
    if (g.M <= 0 || g.N <= 0 || batch <= 0) return kOk;
    if (radon_precision() == 1) {
        const dim3 grid(div_up(g.N, GBN), div_up(g.M, GBM), batch);
        gemm_bf16x3_kernel<<<grid, 256, 0, st>>>(g);
    } else if (g.K < 512) {
        launch_gemm_db<128, 128, 16>(st, g, batch);   // short K: the inverse DFT
    } else {
        launch_gemm_db<64, 64, 32>(st, g, batch);     // long K: the DFT at the bins
    }
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_shot_peak_norm(hipStream_t st, float* a, size_t n, int nshots, float* part)
{
    static_assert(kPeakParts == 64, "the scaling pass combines the partials with one wavefront");
    if (n == 0 || nshots <= 0) return kOk;
    shot_peak_part_kernel<<<dim3(kPeakParts, nshots), 256, 0, st>>>(a, n, part);
    SWD_HIP(hipGetLastError());
    shot_peak_scale_kernel<<<dim3(kPeakParts, nshots), 256, 0, st>>>(a, n, part);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_trace_stats(hipStream_t st, const float* traces, int nt, int ng, int nshots, float* peak, double* par)
{
    if (nt <= 0 || ng <= 0 || nshots <= 0 || (!peak && !par)) return kOk;
    trace_stats_kernel<<<dim3(div_up(ng, 64), nshots), 256, 0, st>>>(traces, nt, ng, peak, par);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_fk_tables(hipStream_t st, const float* W, int nt_eff, int nt, int npair, double* cotab, float* Up,
                     double* EO)
{
    const int npair2 = 2 * npair;
    cot_table_kernel<<<div_up(nt, 256), 256, 0, st>>>(cotab, nt, 2.0 * nt);
    SWD_HIP(hipGetLastError());
    fk_hilbert_table_kernel<<<dim3(div_up(npair2, 64), nt), 64, 0, st>>>(W, nt_eff, nt, npair2, cotab, Up);
    SWD_HIP(hipGetLastError());
    fk_parity_table_kernel<<<div_up(npair2, 64), 64, 0, st>>>(W, nt_eff, npair2, EO);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_fk_window(hipStream_t st, const float* D, const float* H, const double* par, const float* peak,
                     const double* EO, const DispCol* cols, int ncol, int npair, int ng, int Lmax, int nt, float2* X)
{
    if (Lmax > kFkMaxL) { set_error("fk filter windows of %d traces exceed %d", Lmax, kFkMaxL); return kErrArg; }
    const dim3 grid(ncol, div_up(npair, FKK), div_up(Lmax, FKL));
    fk_window_kernel<<<grid, FKL, 0, st>>>(D, H, par, peak, EO, cols, ncol, npair, ng, Lmax, nt, X);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_radon_gather(hipStream_t st, const float* D, const DispCol* cols, int ncol, int npair, int ng, int Lmax,
                        const float* peak, float2* X)
{
    radon_gather_kernel<<<dim3(div_up(ncol, 64), Lmax, npair), 64, 0, st>>>(D, cols, ncol, npair, ng, Lmax, peak, X);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_slant_stack(hipStream_t st, const SlantArgs& a, bool adjoint)
{
    const int M = adjoint ? a.Lmax : a.np;
    const dim3 grid(div_up(a.ncol, GBN), div_up(M, GBM), a.npair);
    if (adjoint) slant_stack_kernel<true><<<grid, 256, 0, st>>>(a);
    else slant_stack_kernel<false><<<grid, 256, 0, st>>>(a);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_radon_abs(hipStream_t st, const float2* MM, int np, int npair, int ncol, float* ml, float* mlr)
{
    radon_abs_kernel<<<dim3(ncol, npair), 256, 0, st>>>(MM, np, npair, ncol, ml, mlr);
    SWD_HIP(hipGetLastError());
    return kOk;
}

template <typename V>
static int pick_launch(hipStream_t st, const V* ml, const V* mlr, const DispCol* cols, int ncol, int np, int npair,
                       int lp, float fg, float lfg, double vmin, int r0, int r1, const double* cr0, double* cr_pre,
                       double* res, int* status)
{
    if (lp == 0 && !mlr) { set_error("an automatic pick start needs |mm|"); return kErrArg; }
    if (np < 1 || np > kPickMaxNp || npair < 1 || npair > kMaxPair) {
        set_error("pick of %d velocities x %d columns outside 1..%d x 1..%d", np, npair, kPickMaxNp, kMaxPair);
        return kErrArg;
    }
    if (np <= 12 * 64)
        pick_kernel<V, 12><<<ncol, 64, 0, st>>>(ml, mlr, cols, np, npair, lp, fg, lfg, vmin, r0, r1, cr0, cr_pre, res,
                                                status);
    else
        pick_kernel<V, kPickMaxNp / 64><<<ncol, 64, 0, st>>>(ml, mlr, cols, np, npair, lp, fg, lfg, vmin, r0, r1, cr0,
                                                             cr_pre, res, status);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_pick(hipStream_t st, const float* ml, const float* mlr, const DispCol* cols, int ncol, int np, int npair,
                int lp, float fg, float lfg, double vmin, int r0, int r1, const double* cr0, double* cr_pre,
                double* res, int* status)
{
    return pick_launch<float>(st, ml, mlr, cols, ncol, np, npair, lp, fg, lfg, vmin, r0, r1, cr0, cr_pre, res, status);
}

int launch_pick(hipStream_t st, const double* ml, const double* mlr, const DispCol* cols, int ncol, int np, int npair,
                int lp, float fg, float lfg, double vmin, int r0, int r1, const double* cr0, double* cr_pre,
                double* res, int* status)
{
    return pick_launch<double>(st, ml, mlr, cols, ncol, np, npair, lp, fg, lfg, vmin, r0, r1, cr0, cr_pre, res,
                               status);
}

int launch_argmax_pick(hipStream_t st, const float* ml, int np, int npair, double vmin, const double* cr0,
                       double* cr_pre)
{
    argmax_pick_kernel<float><<<npair, 64, 0, st>>>(ml, np, npair, vmin, cr0, cr_pre);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_argmax_pick(hipStream_t st, const double* ml, int np, int npair, double vmin, const double* cr0,
                       double* cr_pre)
{
    argmax_pick_kernel<double><<<npair, 64, 0, st>>>(ml, np, npair, vmin, cr0, cr_pre);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_softargmax(hipStream_t st, const float* mlr, const float2* MM, const DispCol* cols, int np, int npair,
                      int ncol, double beta, const double* cr_pre, const double* cr0, float2* GMM)
{
    softargmax_kernel<<<dim3(ncol, npair), 256, 0, st>>>(mlr, MM, cols, np, npair, ncol, beta, cr_pre, cr0, GMM);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_adj_place(hipStream_t st, const float2* GD, const DispCol* cols, int ncol, int npair, int ng, int Lmax,
                     float* Bm)
{
    adj_place_kernel<<<dim3(div_up(npair, 64), Lmax, ncol), 64, 0, st>>>(GD, cols, ncol, npair, ng, Lmax, Bm);
    SWD_HIP(hipGetLastError());
    return kOk;
}

}  // namespace swd
