// Internal declarations shared by the (T) kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scess {

enum Prec : int { P64 = 0, P32 = 1, PB3 = 2 };

// K-dimension chunk (elements) of each precision's W-build kernel; the K
// extents of the packed operands are padded to a multiple of it with zeros.
constexpr int kChunk[3] = {16, 16, 32};
// Bytes of one packed element: FP64 8, FP32 4, BF16X3 4 (hi and lo bf16).
constexpr int kElemBytes[3] = {8, 4, 4};

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Device-resident operands of the W contractions in one precision.
//   X [p][x][y][f]  = (px|yf)         o * v * v * vK
//   T [q][r][z][f]  = t_qr^zf         o * o * v * vK
//   Rn[p][q][x][m]  = -(px|mq)        o * o * v * oK
//   TM[r][y][z][m]  = t_mr^yz         o * v * v * oK
// FP64 and FP32 store plain rows; BF16X3 stores each row as groups of 8 k's:
// 8 hi bf16 followed by 8 lo bf16 (32 bytes per group).
struct Operands {
    void *X = nullptr, *T = nullptr, *Rn = nullptr, *TM = nullptr;
    int64_t vK = 0, oK = 0;   // padded K extents (elements)
    int64_t rbv = 0, rbo = 0; // bytes per packed row of K extent vK / oK
};

// Which source tensor a pack launch reads.
enum PackMode : int { PACK_X = 0, PACK_T = 1, PACK_RN = 2, PACK_TM = 3, PACK_S = 4 };

hipError_t launch_pack(int mode, int prec, const double *src, void *dst, int o, int v, int64_t kpad,
                       hipStream_t s);

struct WBuildArgs {
    const char *X, *T, *Rn, *TM;
    int o, v, vR;
    int64_t rbv, rbo;  // bytes per row (vK, oK rows)
    int nchv, ncho;    // K chunks of a vK row and of an oK row
    int nt;            // tiles per row / column dimension
    int ntile1;        // work-groups of the fixed-a family
    const int *ijk;    // 3 ints per batch entry
    void *W1, *W2;     // accumulator type, [nbatch][v][vR][vR]
    int64_t wstride;   // elements per batch entry
};

// Tile configuration: 2x2 waves, each owning blk x blk MFMA 16x16 blocks.
int wbuild_tile(int prec, int v);  // returns blk for this precision and size
hipError_t launch_wbuild(int prec, int blk, const WBuildArgs &a, int nbatch, hipStream_t s);

struct EpiArgs {
    const void *W1, *W2;  // accumulator type, [nbatch][v][vR][vR]
    int64_t wstride;
    int o, v, vR, nb;     // nb = vR / 8 blocks per dimension
    const int *ijk;       // 3 ints per batch entry
    const int *orbits;    // 3 ints per orbit (A >= B >= C block indices)
    int norb;
    const double *S;      // [o][o][v][v] S[p][q][x][y] = (px|qy)
    const double *t1;     // [o][v]
    const double *e;      // [o+v]
    double *partials;     // [nbatch][norb]
};
hipError_t launch_epilogue(int prec, const EpiArgs &a, int nbatch, hipStream_t s);

// F[ijk_index[b]] = deterministic compensated sum of partials[b][0..norb)
hipError_t launch_reduce_partials(const double *partials, int norb, const int *ijk_index, double *F, int nbatch,
                                  hipStream_t s);
// out[0] = sum_t 2/deg_t F[t] over the unique occupied triples (deterministic, compensated)
hipError_t launch_reduce_energy(const double *F, const int *ijk_all, int nall, double *out, hipStream_t s);

}  // namespace scess
