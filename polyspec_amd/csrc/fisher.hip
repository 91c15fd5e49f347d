// Exact gNL-loc Fisher-matrix derivative for the radial-node optimizer
// (fisher_deriv_gNL_loc, ideal_fisher.pyx:750-787 with _zeta_sum_sym 811-831).
//
//   z^{XY}_l(r, s) = (2l+1) sum_ij X_l^i(r) C^-1_{l,ij} Y_l^j(s),        X, Y in {p, q}
//   S^{XY}(mu; r, s) = sum_l P_l(mu) z^{XY}_l(r, s)
//   D(r, s) = pref w_r w_s sum_mu w_mu [S^PP^3 S^QQ + 3 S^PQ S^QP S^PP^2](mu; r, s)
//
// The l-sum is a GEMM (mu x l) x (l x pairs) for the four tensors PP, PQ, QP, QQ, run on
// v_mfma_f64_16x16x4f64 with the polynomial and the mu-sum fused into its epilogue, so
// S is never written.  Lane layout (verified on gfx950):
//   A[i][k] at lane i + 16 k;  B[k][j] at lane j + 16 k;  D[i][j] at lane j + 16 (i % 4),
//   register i / 4.
// A block = 4 waves = 256 values of mu x one pair tile (r, s0 .. s0 + 15) x 4 tensors; each
// wave holds 4 mu sub-tiles x 4 tensors of 16 x 16 FP64 accumulators.  Only tiles that reach
// the upper triangle s >= r run (D is symmetric); the last pass adds the mu-block partials
// in a fixed order, scales by pref w_r w_s and mirrors.
#include <algorithm>

#include "polyspec_hip.h"
#include "ps_common.h"

namespace ps {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int FD_LCH = 16;   // l per staged chunk (4 MFMA k-steps)
constexpr int FD_MUB = 256;  // mu per block
constexpr int FD_PT = 16;    // pairs per tile

struct FdDims {
    int nl, npol, nr, nmu, lmin;
    int NLP, NRP, NMUP;  // padded l (multiple of 16), s (multiple of 16), mu (multiple of 256)
    long ntile;
};

// Z[X][l][r][s] for X = PP, PQ, QP, QQ; zero for l >= nl or s >= nr.
__global__ void k_fd_zeta(FdDims d, const double* __restrict__ P, const double* __restrict__ Q,
                          const double* __restrict__ Cinv, double* __restrict__ Z) {
    const long n = (long)d.NLP * d.nr * d.NRP;
    const long zs = n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int s = (int)(e % d.NRP);
        const long t = e / d.NRP;
        const int r = (int)(t % d.nr);
        const int l = (int)(t / d.nr);
        double pp = 0.0, pq = 0.0, qp = 0.0, qq = 0.0;
        if (l < d.nl && s < d.nr) {
            for (int i = 0; i < d.npol; i++) {
                const double pr = P[((long)l * d.npol + i) * d.nr + r], qr = Q[((long)l * d.npol + i) * d.nr + r];
                for (int j = 0; j < d.npol; j++) {
                    const double c = Cinv[((long)i * d.npol + j) * d.nl + l];
                    const double ps_ = P[((long)l * d.npol + j) * d.nr + s], qs = Q[((long)l * d.npol + j) * d.nr + s];
                    pp += pr * c * ps_;
                    pq += pr * c * qs;
                    qp += qr * c * ps_;
                    qq += qr * c * qs;
                }
            }
            const double tw = 2.0 * (d.lmin + l) + 1.0;
            pp *= tw;
            pq *= tw;
            qp *= tw;
            qq *= tw;
        }
        Z[e] = pp;
        Z[zs + e] = pq;
        Z[2 * zs + e] = qp;
        Z[3 * zs + e] = qq;
    }
}

// legsT[l][mu] = legs[mu][l] (zero padded), wmu padded with zeros
__global__ void k_fd_legs(FdDims d, const double* __restrict__ legs, const double* __restrict__ wmu,
                          double* __restrict__ legsT, double* __restrict__ wmup) {
    const long n = (long)d.NLP * d.NMUP;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int mu = (int)(e % d.NMUP), l = (int)(e / d.NMUP);
        legsT[e] = (l < d.nl && mu < d.nmu) ? legs[(long)mu * d.nl + l] : 0.0;
        if (l == 0) wmup[mu] = mu < d.nmu ? wmu[mu] : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_fd_gemm(FdDims d, const double* __restrict__ Z,
                                                 const double* __restrict__ legsT, const double* __restrict__ wmup,
                                                 const int2* __restrict__ tiles, double* __restrict__ part) {
    __shared__ __align__(32) double sB[FD_LCH * 4 * FD_PT];  // [l][X][j]
    __shared__ double red[4][FD_PT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, g = lane >> 4;
    // mu blocks of one pair tile are consecutive logical blocks on one XCD, so the tile's
    // zeta columns come from HBM once and are re-read from that XCD's L2
    const int nmb = d.NMUP / FD_MUB;
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const long tile = lb / nmb;
    const int mb = lb % nmb;
    const int2 rt = tiles[tile];
    const int r = rt.x, s0 = rt.y * FD_PT;
    const int mu0 = mb * FD_MUB + wave * 64;
    const bool live = mu0 < d.nmu;  // waves past the last node only help staging
    const long zs = (long)d.NLP * d.nr * d.NRP;
    // staging: thread t loads 4 consecutive s of (l = t / 16, X = (t / 4) % 4)
    const int sl = tid >> 4, sX = (tid >> 2) & 3, sj = (tid & 3) * 4;
    const double* zsrc = Z + sX * zs + (long)r * d.NRP + s0 + sj;
    const long lstride = (long)d.nr * d.NRP;
    double4 pb = *reinterpret_cast<const double4*>(zsrc + (long)sl * lstride);
    double pa[4][4];  // [u][k-step] A operand for this lane: legs[mu0 + 16u + j][l0 + 4 ks + g]
    auto fetchA = [&](int l0) {
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int u = 0; u < 4; u++) pa[u][ks] = legsT[(long)(l0 + 4 * ks + g) * d.NMUP + mu0 + 16 * u + j];
    };
    if (live) fetchA(0);
    v4d acc[4][4];  // [X][u]
#pragma unroll
    for (int X = 0; X < 4; X++)
#pragma unroll
        for (int u = 0; u < 4; u++) acc[X][u] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int l0 = 0; l0 < d.NLP; l0 += FD_LCH) {
        __syncthreads();
        *reinterpret_cast<double4*>(&sB[(sl * 4 + sX) * FD_PT + sj]) = pb;
        double a[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int ks = 0; ks < 4; ks++) a[u][ks] = pa[u][ks];
        __syncthreads();
        if (l0 + FD_LCH < d.NLP) {
            pb = *reinterpret_cast<const double4*>(zsrc + (long)(l0 + FD_LCH + sl) * lstride);
            if (live) fetchA(l0 + FD_LCH);
        }
        if (!live) continue;
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            double b[4];
#pragma unroll
            for (int X = 0; X < 4; X++) b[X] = sB[((4 * ks + g) * 4 + X) * FD_PT + j];
#pragma unroll
            for (int X = 0; X < 4; X++)
#pragma unroll
                for (int u = 0; u < 4; u++)
                    acc[X][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][ks], b[X], acc[X][u], 0, 0, 0);
        }
    }
    // epilogue: lane (j, g), register v of sub-tile u holds mu = mu0 + 16 u + 4 v + g, pair j
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double AA = acc[0][u][v], AB = acc[1][u][v], BA = acc[2][u][v], BB = acc[3][u][v];
            const double A2 = AA * AA;
            sum += wmup[mu0 + 16 * u + 4 * v + g] * (A2 * AA * BB + 3.0 * AB * BA * A2);
        }
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    if (g == 0) red[wave][j] = sum;
    __syncthreads();
    if (tid < FD_PT)
        part[((long)mb * d.ntile + tile) * FD_PT + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// First tile of row r: sum over r' < r of the tiles of row r', (NRP/16 - r'/16).
__host__ __device__ inline long fd_row_tile0(int r, int K) {
    const long q = r / FD_PT, rem = r % FD_PT;
    return (long)r * K - (FD_PT * q * (q - 1) / 2 + rem * q);
}

// tile list: row r owns (r, s-tile) for s-tiles r/16 .. NRP/16 - 1
__global__ void k_fd_tiles(FdDims d, int2* __restrict__ tiles) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.nr) return;
    const int K = d.NRP / FD_PT;
    long t = fd_row_tile0(r, K);
    for (int st = r / FD_PT; st < K; st++) tiles[t++] = int2{r, st};
}

// D[r][s] = D[s][r] = pref w_r w_s sum_mb part[mb][tile(r, s / 16)][s % 16] for s >= r
__global__ void k_fd_finish(FdDims d, const double* __restrict__ part, const double* __restrict__ w, double pref,
                            double* __restrict__ out) {
    const long n = (long)d.nr * d.nr;
    const int nmb = d.NMUP / FD_MUB;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e / d.nr), s = (int)(e % d.nr);
        if (s < r) continue;
        const long tile = fd_row_tile0(r, d.NRP / FD_PT) + (s / FD_PT - r / FD_PT);
        double v = 0.0;
        for (int mb = 0; mb < nmb; mb++) v += part[((long)mb * d.ntile + tile) * FD_PT + s % FD_PT];
        v *= pref * w[r] * w[s];
        out[(long)r * d.nr + s] = v;
        out[(long)s * d.nr + r] = v;
    }
}

static FdDims fd_dims(int nl, int npol, int nr, int nmu, int lmin) {
    FdDims d;
    d.nl = nl;
    d.npol = npol;
    d.nr = nr;
    d.nmu = nmu;
    d.lmin = lmin;
    d.NLP = (nl + FD_LCH - 1) / FD_LCH * FD_LCH;
    d.NRP = (nr + FD_PT - 1) / FD_PT * FD_PT;
    d.NMUP = (nmu + FD_MUB - 1) / FD_MUB * FD_MUB;
    long nt = 0;
    for (int r = 0; r < nr; r++) nt += d.NRP / FD_PT - r / FD_PT;
    d.ntile = nt;
    return d;
}

struct FdWork {
    size_t z, legs, wmu, tiles, part, total;
};

static FdWork fd_work(const FdDims& d) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    FdWork w;
    w.z = 0;
    size_t o = al((size_t)4 * d.NLP * d.nr * d.NRP * sizeof(double));
    w.legs = o;
    o += al((size_t)d.NLP * d.NMUP * sizeof(double));
    w.wmu = o;
    o += al((size_t)d.NMUP * sizeof(double));
    w.tiles = o;
    o += al((size_t)d.ntile * sizeof(int2));
    w.part = o;
    o += al((size_t)(d.NMUP / FD_MUB) * d.ntile * FD_PT * sizeof(double));
    w.total = o;
    return w;
}

}  // namespace ps

using namespace ps;

extern "C" {

long long ps_fisher_deriv_work_bytes(int nl, int npol, int nr, int nmu) {
    if (nl < 1 || npol < 1 || nr < 1 || nmu < 1) return -1;
    return (long long)fd_work(fd_dims(nl, npol, nr, nmu, 0)).total;
}

int ps_fisher_deriv_gnl_loc(const double* P, const double* Q, int nl, int npol, int nr, const double* Cinv,
                            const double* legs, const double* w_mu, int nmu, const double* weights, int lmin,
                            void* work, long long work_bytes, double* out, void* stream) {
    PS_REQUIRE(P && Q && Cinv && legs && w_mu && weights && work && out, "ps_fisher_deriv_gnl_loc: NULL argument");
    PS_REQUIRE(nl >= 1 && npol >= 1 && npol <= 3 && nr >= 1 && nmu >= 1 && lmin >= 0,
               "ps_fisher_deriv_gnl_loc: bad sizes");
    const FdDims d = fd_dims(nl, npol, nr, nmu, lmin);
    const FdWork wk = fd_work(d);
    PS_REQUIRE(work_bytes >= (long long)wk.total, "ps_fisher_deriv_gnl_loc: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    char* base = (char*)work;
    double* Z = (double*)(base + wk.z);
    double* legsT = (double*)(base + wk.legs);
    double* wmup = (double*)(base + wk.wmu);
    int2* tiles = (int2*)(base + wk.tiles);
    double* part = (double*)(base + wk.part);
    hipLaunchKernelGGL(k_fd_tiles, dim3((nr + 255) / 256), dim3(256), 0, st, d, tiles);
    PS_LAUNCH_CHECK();
    const long nz = (long)d.NLP * d.nr * d.NRP;
    hipLaunchKernelGGL(k_fd_zeta, dim3((unsigned)std::min<long>((nz + 255) / 256, 65536)), dim3(256), 0, st, d, P,
                       Q, Cinv, Z);
    PS_LAUNCH_CHECK();
    const long nlg = (long)d.NLP * d.NMUP;
    hipLaunchKernelGGL(k_fd_legs, dim3((unsigned)std::min<long>((nlg + 255) / 256, 65536)), dim3(256), 0, st, d, legs,
                       w_mu, legsT, wmup);
    PS_LAUNCH_CHECK();
    PS_REQUIRE(d.ntile * (d.NMUP / FD_MUB) < 2147483647L, "ps_fisher_deriv_gnl_loc: too many radial nodes");
    hipLaunchKernelGGL(k_fd_gemm, dim3((unsigned)(d.ntile * (d.NMUP / FD_MUB))), dim3(256), 0, st, d, Z, legsT, wmup,
                       tiles, part);
    PS_LAUNCH_CHECK();
    const long nn = (long)nr * nr;
    const double pref = 6.0 / (16.0 * M_PI * M_PI) * (81.0 / 625.0) / 2.0;
    hipLaunchKernelGGL(k_fd_finish, dim3((unsigned)std::min<long>((nn + 255) / 256, 65536)), dim3(256), 0, st, d, part,
                       weights, pref, out);
    PS_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
