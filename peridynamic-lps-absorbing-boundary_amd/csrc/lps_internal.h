// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Device data layout (all arrays in the permuted node order):
//   [0, n_int)       non-ABC nodes in Morton order of their lattice position
//   [n_int, n)       active ABC nodes (bctypes == 1), one per ABC row, in row order
// Fields are SoA (x and y components in separate arrays) so that a wavefront's 64 lanes load
// 64 consecutive doubles. Horizons are kept twice: as CSR (self first, distance-sorted) for the
// setup passes and the export, and as SELL-64 (self dropped, slice = one wavefront of rows,
// column-major inside a slice) for the per-step passes, so that every index load is one
// coalesced 256-byte line per wavefront.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lps {

constexpr int kSlice = 64;      // SELL slice height = wavefront size
constexpr int kBlock = 256;     // threads per block of the node passes (4 wavefronts)
constexpr int kMaxHorizon = 128;  // MAX_HORIZON_LENGTH, src/main.cpp:35
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kMaxGeoK = 6;       // largest lattice-offset radius of the fast path's LDS table

// Bond geometry of one lattice offset (fast path): unit vector, |xi| beta dV, beta dV.
struct Geo {
  double hx, hy, g1, g2;
};

// Everything a step kernel reads; device pointers.
struct StepArgs {
  int32_t n, n_int, n_abc, M;      // nodes, non-ABC nodes, ABC rows, cloud width
  double dt, rho, lam, mu, kappa, delta, dx, dy;
  // geometry
  const double *x, *y;           // coordinates
  const int32_t *lat;            // packed lattice coords (ix << 16 | iy), fast path
  const Geo *geo;                // (2K+1)^2 offset table, fast path
  int32_t geo_k;                 // K: offsets in [-K, K]
  // horizons
  const int64_t *row_ptr;        // CSR, self first
  const int32_t *col;
  const int64_t *sell_off;       // per slice, in elements
  const int32_t *sell_w;         // per slice width (self excluded)
  const int32_t *sell_col;       // -1 padded
  // per node constants
  const double *m;               // weighted volume m_i
  const double *bcoef;           // 8 mu / m_i, or -1 when m_i <= 0 (fast path)
  const double *c2m;             // 2 / m_i, or 0 when m_i <= 0 (fast path)
  const int8_t *is_bc;           // bctypes[2i] == 1 (IDW neighbour filter)
  // ABC rows
  const int32_t *cloud;          // n_abc x M, permuted ids, -1 = skip
  const double *gu, *gv;         // n_abc x 2 x 4M
  const double *gth;             // n_abc x 2M
  int32_t theta_method;
  int32_t abc_on;
};

// State buffers for one step (SoA).
struct StepState {
  const double *px, *py;         // prediction U^{n+1} before the ABC overwrite
  const double *ux, *uy, *uxm, *uym;   // U^n, U^{n-1}
  const double *vx, *vy, *vxm, *vym;   // V^n, V^{n-1}
  const double *ax, *ay;         // A^n
  double *theta;                 // theta^{n+1} (written by the dilatation pass)
  double *acoef;                 // 2 (lambda - mu) theta_i / m_i (fast path, written with theta)
  const double *thn, *thnm1;     // theta history
  double *ux1, *uy1, *vx1, *vy1, *ax1, *ay1;  // U^{n+1}, V^{n+1}, A^{n+1}
  double *pnx, *pny;             // next prediction U^{n+2} (fused predictor)
};

// Per-block partial sums of the output block.
struct DiagPartial {
  double energy, ek, ep, px, py, max_u, max_v, max_a, th_max, th_min, ub, vb;
  int64_t max_u_orig, max_v_orig;
};

// Kernel launchers (lps_kernels.hip).
hipError_t launch_predict(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                          const double *vy, const double *ax, const double *ay, double *px,
                          double *py, hipStream_t s);
hipError_t launch_theta(const StepArgs &a, const StepState &st, bool exact, hipStream_t s);
hipError_t launch_idw(const StepArgs &a, const StepState &st, bool exact, hipStream_t s);
hipError_t launch_force(const StepArgs &a, const StepState &st, bool exact, hipStream_t s);
hipError_t launch_theta_field(const StepArgs &a, const double *ux, const double *uy, double *theta,
                              hipStream_t s);
hipError_t launch_weighted_volume(const StepArgs &a, double *m, hipStream_t s);
hipError_t launch_diag(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                       const double *vy, const double *ax, const double *ay, const double *theta,
                       const int32_t *perm, double exclude_r, bool exact, DiagPartial *partials,
                       int32_t n_partials, DiagPartial *out, hipStream_t s);
int32_t diag_partials(int32_t n);

// Horizon build (lps_neighbors.hip). Builds perm (new -> original), CSR and SELL on the device.
struct HorizonBuild {
  int64_t nnz = 0;                 // CSR entries including self
  int32_t *perm = nullptr;         // n
  int32_t *iperm = nullptr;        // n
  int64_t *row_ptr = nullptr;      // n + 1
  int32_t *col = nullptr;          // nnz
  int64_t *sell_off = nullptr;     // slices + 1
  int32_t *sell_w = nullptr;       // slices
  int32_t *sell_col = nullptr;
  int64_t sell_len = 0;
  int32_t max_len = 0;
};
// xo, yo: original coordinates on the device; seg: 1 for nodes of the ABC segment (original order),
// seg_rank: their rank inside the segment (ABC row). Returns hipSuccess or an error; *overflow is
// set when a horizon exceeds max_len.
hipError_t build_horizons(int32_t n, const double *xo, const double *yo, const int8_t *seg,
                          const int32_t *seg_rank, double r, int32_t max_len, HorizonBuild *hb,
                          bool *overflow, hipStream_t s);
void free_horizons(HorizonBuild *hb);

}  // namespace lps
