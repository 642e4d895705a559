// FP64 GEMM on the matrix cores for the multiscale hat-matrix update
// (GWRMultiscale.cpp:296-301: R_i <- S_i (R_i - S0) + S_i): C += A B for
// n x n column-major matrices.
//
// One workgroup (8 waves, 2 x 4) computes a 128 x 128 tile of C; each wave a
// 64 x 32 part as 4 x 2 v_mfma_f64_16x16x4f64 tiles.  A and B stream through
// LDS in K slices of 16, double buffered (66 KB), so two workgroups share a
// CU and four waves a SIMD: while one waits on LDS or a barrier the others
// keep the matrix core busy.  The next slice is loaded into registers while
// the current one is multiplied and written to the other buffer afterwards.
// Slices are stored k-major with a row stride of 129 doubles, so the 16
// consecutive rows (columns) a quarter-wave reads lie in distinct banks.
// Measured at n = 4000: 2.58 ms, 63% of the 78.6 TFLOP/s FP64 peak (one wave
// per SIMD with 64 x 64 wave tiles and 32-deep slices: 2.94 ms).
//
// Fragment order: the MFMA computes the transposed tile, D = B^T A^T, whose
// 16 x 16 layout puts C's rows along the lanes (lane % 16), so the epilogue's
// read-modify-write of column-major C touches 16 consecutive doubles per
// quarter-wave.
//
// Tile order: workgroup b runs on XCD b % 8; each XCD takes a contiguous range
// of tiles, visited in groups of 4 tile rows, so the workgroups an XCD holds
// at once share a few row panels of A and column panels of B in its L2.
#include "gwm_contract.h"

namespace gwm {

namespace {

constexpr int kTile = 128;     // C tile edge
constexpr int kLdp = 129;      // LDS row stride, doubles
constexpr int kGroup = 4;      // tile rows per group of the visiting order
constexpr int kXcds = 8;
constexpr int kWM = 2, kWN = 4;   // waves along rows / columns
constexpr int kKT = 16;           // K slice

__global__ __launch_bounds__(64 * kWM * kWN) __attribute__((amdgpu_waves_per_eu(4, 4))) void dgemm_acc_kernel(int n, const double* __restrict__ A,
                                                                 const double* __restrict__ B,
                                                                 double* __restrict__ C, int tiles, int per_xcd) {
    constexpr int NT = 64 * kWM * kWN;           // threads
    constexpr int TI = kTile / kWM / 16, TJ = kTile / kWN / 16;   // MFMA tiles per wave
    constexpr int PA = kTile * kKT / NT;         // A (and B) elements staged per thread
    __shared__ double As[2][kKT][kLdp];
    __shared__ double Bs[2][kKT][kLdp];
    const int q = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    if (q >= tiles * tiles) return;
    const int grp = q / (kGroup * tiles);
    const int first = grp * kGroup;
    const int gm = min(tiles - first, kGroup);
    const int rem = q % (kGroup * tiles);
    const int m0 = (first + rem % gm) * kTile;
    const int n0 = (rem / gm) * kTile;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wr = (wave % kWM) * (kTile / kWM), wc = (wave / kWM) * (kTile / kWN);

    // staging: A rows m0 + t % 128, k = t / 128 + (NT / 128) p; B k = t % kKT,
    // columns n0 + t / kKT + (NT / kKT) p.  Loads take clamped in-range
    // addresses; out-of-range elements are zeroed when the registers are
    // written to LDS, so a load is consumed only one slice later.
    double ra[PA], rb[PA];
    const int arow = m0 + (t & 127), ak = t >> 7;
    const int bk = t % kKT, bcol = n0 + t / kKT;
    const int arow_c = min(arow, n - 1);
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PA; ++p) ra[p] = A[(size_t)min(k0 + ak + (NT / 128) * p, n - 1) * n + arow_c];
#pragma unroll
        for (int p = 0; p < PA; ++p) rb[p] = B[(size_t)min(bcol + (NT / kKT) * p, n - 1) * n + min(k0 + bk, n - 1)];
    };
    auto store = [&](int buf, int k0) {
#pragma unroll
        for (int p = 0; p < PA; ++p)
            As[buf][ak + (NT / 128) * p][t & 127] = (arow < n && k0 + ak + (NT / 128) * p < n) ? ra[p] : 0.0;
#pragma unroll
        for (int p = 0; p < PA; ++p)
            Bs[buf][bk][t / kKT + (NT / kKT) * p] = (bcol + (NT / kKT) * p < n && k0 + bk < n) ? rb[p] : 0.0;
    };

    v4d acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    const int nk = (n + kKT - 1) / kKT;
    load(0);
    store(0, 0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        load((kt + 1) * kKT);   // past the end: clamped, never stored
#pragma unroll
        for (int ks = 0; ks < kKT / 4; ++ks) {
            const int k = 4 * ks + lk;
            double a[TI], b[TJ];
#pragma unroll
            for (int i = 0; i < TI; ++i) a[i] = As[buf][k][wr + 16 * i + li];
#pragma unroll
            for (int j = 0; j < TJ; ++j) b[j] = Bs[buf][k][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) store(buf ^ 1, (kt + 1) * kKT);
        __syncthreads();
    }
    // acc[i][j][r]: C row m0 + wr + 16 i + li, column n0 + wc + 16 j + lk + 4 r
#pragma unroll
    for (int i = 0; i < TI; ++i) {
        const int row = m0 + wr + 16 * i + li;
        if (row >= n) continue;
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = n0 + wc + 16 * j + lk + 4 * r;
                if (col < n) C[(size_t)col * n + row] += acc[i][j][r];
            }
    }
}

}  // namespace

hipError_t launch_dgemm_acc(int n, const double* A, const double* B, double* C, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int tiles = (n + kTile - 1) / kTile;
    const int per_xcd = (tiles * tiles + kXcds - 1) / kXcds;
    hipLaunchKernelGGL(dgemm_acc_kernel, dim3(per_xcd * kXcds), dim3(64 * kWM * kWN), 0, s, n, A, B, C, tiles, per_xcd);
    return hipGetLastError();
}

}  // namespace gwm
