// tables.hpp -- host-side FP64 construction of the constant tables the
// kernels read (built once per bandwidth / voxel count, stored as FP32).
#pragma once

#include <cstdint>
#include <vector>

namespace phz {

// Half-open index run [start, next start) fused with one mode:
// 0 = max-energy channel, 1 = mean over all channels, 2 = zero.
struct FusionRuns {
  static constexpr int kMax = 32;
  int count = 0;
  uint32_t start[kMax];
  uint32_t mode[kMax];
};

struct S2Tables {
  int bw = 0, n = 0;
  std::vector<float> grid;        // n*n*4: unit query vectors (x,y,z,0), spherical-sampler.cc:29-64
  std::vector<float> theta;       // n: colatitude of each grid row (exact, for NN bounds)
  std::vector<float> legendre;    // [m][i][l-m]: w_i * P~_l^m(cos th_i)
  std::vector<int> leg_off;       // per m offset into legendre
  std::vector<float> wigner;      // blocks (a<=b): [l-b][k] d~^l_{a,b}(beta_k)
  std::vector<int64_t> wig_off;   // per block id (b*(b+1)/2 + a)
  std::vector<float> wig_norm;    // 2 pi sqrt(2/(2l+1)), so3_correlate_fftw.c:120
  std::vector<float> tw_n;        // 2n floats: exp(-2 pi i e / n)
  FusionRuns runs;                // coefficient fusion in unshifted seanindex order
};

struct VoxelTables {
  int N = 0;
  float lo = 0, hi = 0;
  std::vector<float> edges;   // LinSpaced(N, lo, hi)
  std::vector<float> tw_N;    // 2N floats
  FusionRuns runs;            // over the first N^2(N/2+1) unshifted entries
};

void build_s2_tables(int bw, S2Tables* t);
void build_voxel_tables(int N, float lo, float hi, VoxelTables* t);
// seanindex, s2_primitive.c:300-313
int seanindex(int m, int l, int bw);
// run-length fusion modes of an n-entry pyramid (5 levels, divider 4),
// expressed for unshifted index q; only q < limit is described.
bool fusion_runs(uint64_t n, uint64_t limit, FusionRuns* out);
std::vector<float> twiddles(int L);
// PyramidStruct bounds (pyramid-struct.cc:17-29)
void pyramid_bounds(uint32_t n, int levels, float div, uint32_t* coeffs, uint32_t* lower,
                    uint32_t* upper);

}  // namespace phz
