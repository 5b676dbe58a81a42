// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Device data layout. Every per-node array is indexed by "slot":
//   tile layout (fast path, lattice inputs): the lattice is cut into 16 x 16 tiles; each tile that
//     holds a node owns 256 consecutive slots (x-major, y fastest inside the tile); lattice points
//     without a node are holes (kind 0). ABC nodes stay in place.
//   Morton layout (exact path, or nodes not on a lattice): non-ABC nodes in Morton order of their
//     position, then the ABC nodes in ABC-row order; no holes.
// kind[slot]: 0 hole, 1 node, 2 node of an active ABC row; abc_slot[row] gives the slot of a row.
// Fields are SoA (x and y components in separate arrays), so a wavefront's 64 lanes load 64
// consecutive doubles. Horizons exist as CSR (self first, distance-sorted; the setup passes, the
// ABC rows and the export read it) and as SELL-64 (self dropped, one slice = one wavefront of
// rows, column-major inside a slice) for the exact per-step passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace lps {

constexpr int kSlice = 64;        // SELL slice height = wavefront size
constexpr int kBlock = 256;       // threads per block of the node passes (4 wavefronts)
constexpr int kWavesPerBlock = kBlock / 64;

// Blocks are spread over the 8 XCDs round-robin (block b runs on XCD b % 8). Work item
// (b % 8) * per_xcd + b / 8 of n goes to block b, so each XCD takes a contiguous range and
// neighbouring items share its L2; launch ceil(n / 8) * 8 blocks (items past n: the block exits).
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t n) {
  const int32_t per_xcd = (n + 7) / 8;
  return (b % 8) * per_xcd + b / 8;
}
constexpr int kMaxHorizon = 128;  // MAX_HORIZON_LENGTH, src/main.cpp:35
constexpr int kTile = 16;         // tile edge (lattice points)
constexpr int kMaxHalo = 4;       // largest stencil reach of the tile path (lattice cells)
constexpr int kMaxWin = kTile + 2 * kMaxHalo;  // 24

// Tile-path geometry: the stencil (bond offsets of a full horizon, in summation order) and the
// tile map. Bond k of slot (lx, ly) reads window cell c + delta[k].
struct TileArgs {
  int32_t n_tiles, ntx, nty;       // active tiles; tile grid
  int32_t halo, win;               // stencil reach K, window edge 16 + 2K
  const int32_t *tile_xy;          // n_tiles: tx << 16 | ty
  const int32_t *tile_map;         // ntx * nty -> tile index or -1
  const int32_t *tile_nbr;         // n_tiles x 9: the 3 x 3 tiles around a tile (-1 absent)
  const int8_t *tile_full;         // n_tiles: every window cell holds a node with m == m_ref
  const int32_t *full_list;        // the full tiles, in tile order
  const int32_t *part_list;        // the other tiles
  int32_t n_full, n_part;
  int32_t nt;                      // stencil bonds
  const int32_t *delta;            // nt, window-cell offset ox * win + oy
  const double *hx, *hy, *g1, *g2; // nt: unit vector, |xi| beta dV, beta dV
  const double *xi;                // nt: |xi|
  // The reference stencil (dx == dy, delta = 4 dx: the 68 offsets 0 < ox^2 + oy^2 <= 20) has 12
  // distance classes; with c = dx / |xi|, class constants c1 = |xi| beta dV c and
  // c2 = beta dV c^2 turn the bond terms into integer multiples (xi_hat = (ox, oy) c).
  // Node kernel of each pass: kGeneric (any stencil, table-driven), kUnrolled (68 bonds, one node
  // per thread), kBlocked (68 bonds, four nodes per thread on the full tiles).
  int32_t theta_kernel, force_kernel;
  double c1[12], c2[12];
  // Full tiles, where every window cell holds a node with m == m_ref (C = 2 / m_ref): class
  // constants folded with C, for the class representative (a, b), a >= b >= 0, a^2 + b^2 = r^2:
  //   wt[cls][0 / 1]     = C c1 {a, b}               (dilatation; force, theta_j terms)
  //   wb[cls][0 / 1 / 2] = 8 C c2 {a^2, a b, b^2}    (force, bond-extension terms)
  // The force of an instance uses theta scaled by InstParams.th_scale = (lambda - mu) / mu and
  // the sum scaled by InstParams.f_scale = mu, so the tables hold no material and an instance's
  // arithmetic is the same alone or in a batch. These tables are kernel arguments; the kernel
  // argument segment is kept small (1.7 KB instead of 1.4 KB made the force kernel 35% slower).
  double wt[12][2], wb[12][3];
  double wbsum[2];                 // sum over the stencil of wb of o_x^2 and of o_y^2
  // register-blocked kernels: nodes per lane, 4 (a wavefront per tile) or 2 (two per tile)
  int32_t theta_nr, force_nr;
  const int32_t *tile_inst;        // n_tiles: the instance (domain) a tile belongs to
};

// Per instance (one domain of a batch; a single problem is a batch of one): time step, material,
// and the factors of the full tiles' force with the tables of TileArgs:
// th_scale = (lambda - mu) / mu, f_scale = mu.
struct InstParams {
  double dt, rho, lam, mu, kappa, exclude_r;
  double th_scale, f_scale;
};

enum NodeKernel : int32_t { kGeneric = 0, kUnrolled = 1, kBlocked = 2 };

// Distance classes of the 68-bond stencil: ox^2 + oy^2.
constexpr int kStencilR2[12] = {1, 2, 4, 5, 8, 9, 10, 13, 16, 17, 18, 20};

// Everything a step kernel reads; device pointers.
struct StepArgs {
  int32_t n;                       // slots
  int32_t n_abc, M;                // ABC rows, cloud width
  double dt, rho, lam, mu, kappa, delta, dx, dy;
  const double *x, *y;             // coordinates (0 at holes)
  const int8_t *kind;              // 0 hole, 1 node, 2 ABC-row node
  const int32_t *abc_slot;         // n_abc
  // horizons
  const int64_t *row_ptr;          // CSR, self first (holes: empty)
  const int32_t *col;
  const int64_t *sell_off;         // per slice, in elements
  const int32_t *sell_w;           // per slice width (self excluded)
  const int32_t *sell_col;         // -1 padded
  // per node constants
  const double *m;                 // weighted volume m_i
  const double *c2m;               // 2 / m_i, or 0 when m_i <= 0
  const int8_t *is_bc;             // bctypes[2i] == 1 (IDW neighbour filter)
  // ABC rows, ragged: row r has the L = abc_off[r + 1] - abc_off[r] cloud members from
  // o = abc_off[r] on (at most M; the tile path drops trailing members that add nothing)
  const int32_t *abc_off;          // n_abc + 1
  const int32_t *cloud;            // member slots, -1 = skip
  // Operator blocks: a row's G_u / G_v / G_theta coefficients are those of an operator block of L
  // members at member offset g = abc_gofs[r] (the x block, 4 L doubles at 8 g, then the y block;
  // G_theta 2 per member at 2 g), its member jj being the block's member abc_pmap[o + jj], mapped
  // by the lattice transform abc_xf[r] (abc_uv_coeffs). A row with operators of its own has its own
  // block, the identity map and transform 0; rows that are images of one row under the square
  // lattice's symmetries share that row's block.
  const int32_t *abc_gofs;         // n_abc
  const int8_t *abc_xf;            // n_abc: bit 0 swap x and y, bit 1 negate the cross terms
  const uint8_t *abc_pmap;         // per member
  const double *gu, *gv;
  const double *gth;
  int32_t theta_method;
  const InstParams *inst;          // per instance; the fast-path kernels read it through
                                   // tile.tile_inst, the exact path uses the scalars above
  TileArgs tile;
};

// Where ABC row `row` lies in the ragged arrays: its first member and its member count.
__device__ __forceinline__ void abc_row_span(const StepArgs &a, int32_t row, int64_t &o, int32_t &L) {
  const int32_t o0 = a.abc_off[row];
  o = o0;
  L = a.abc_off[row + 1] - o0;
}

// A 2 x 2 x 2 block of G_u or G_v (b0: into x, b1: into y; each (U^n_x, U^n_y, U^{n-1}_x,
// U^{n-1}_y)) under the lattice transform T of code xf: T B T' is B with x and y swapped (bit 0)
// and the cross terms (x into y, y into x) negated (bit 1), since T is a signed permutation. Exact.
__device__ __forceinline__ void lattice_map(int xf, double4 b0, double4 b1, double4 &o0, double4 &o1) {
  if (xf & 1) {
    o0 = make_double4(b1.y, b1.x, b1.w, b1.z);
    o1 = make_double4(b0.y, b0.x, b0.w, b0.z);
  } else {
    o0 = b0;
    o1 = b1;
  }
  if (xf & 2) {
    o0.y = -o0.y; o0.w = -o0.w;
    o1.x = -o1.x; o1.z = -o1.z;
  }
}

// G_u (g0, g1) and G_v (h0, h1) coefficients of member jj of a row (members from o, L of them,
// block at g, transform xf).
__device__ __forceinline__ void abc_uv_coeffs(const StepArgs &a, int64_t o, int32_t L, int64_t g,
                                              int xf, int jj, double4 &g0, double4 &g1, double4 &h0,
                                              double4 &h1) {
  const int m = a.abc_pmap[o + jj];
  const double *bu = a.gu + 8 * g + 4 * m, *bv = a.gv + 8 * g + 4 * m;
  lattice_map(xf, *(const double4 *)bu, *(const double4 *)(bu + 4 * L), g0, g1);
  lattice_map(xf, *(const double4 *)bv, *(const double4 *)(bv + 4 * L), h0, h1);
}

// State buffers for one step (SoA).
struct StepState {
  const double *px, *py;           // prediction U^{n+1} before the ABC overwrite
  const double *ux, *uy, *uxm, *uym;   // U^n, U^{n-1}
  const double *vx, *vy, *vxm, *vym;   // V^n, V^{n-1}
  const double *ax, *ay;           // A^n
  double *theta;                   // theta^{n+1} (written by the dilatation pass)
  const double *thn, *thnm1;       // theta history
  double *ux1, *uy1, *vx1, *vy1, *ax1, *ay1;  // U^{n+1}, V^{n+1}, A^{n+1}
  double *pnx, *pny;               // next prediction U^{n+2} (fused predictor)
};

// Fast path: U^n and U^{n-1} at the interior nodes (kind 1). The corrector leaves U^{n+1} equal
// to the prediction there, so those values stay only in the ring of earlier predictions; the
// ABC nodes' values are in StepState.ux / uxm.
struct RingU {
  const double *ux, *uy, *uxm, *uym;
};

// Fast path, ABC rows: the state of the clouds' members packed per member at the step's start
// (launch_abc_pack), so a member's G_u / G_v terms gather two 32-byte records (and G_theta one
// 16-byte record) instead of eight (two) values from separate arrays. The records of step n's
// U^n, V^n are those of step n + 1's U^{n}, V^{n} (two buffers by step parity), so a step packs
// one level.
struct AbcPack {
  // per row, stride entries (the widest row's length): the row's cloud entries in the order of its
  // operator block (entry jj of row r at r * stride + abc_pmap[abc_off[r] + jj]), each its member's
  // record, -1 for a pad and past the row's end. The fixed stride lets a row kernel load the
  // entries before it knows where the row's ragged data start.
  const int32_t *idx;
  const double4 *cur;  // per member: (U^n_x, U^n_y, V^n_x, V^n_y)
  const double4 *prev; // per member: (U^{n-1}_x, U^{n-1}_y, V^{n-1}_x, V^{n-1}_y), the previous
                       // step's cur
  const double2 *th;   // per member: (theta^n, theta^{n-1}) (G_theta only)
  int32_t stride;
};

// Per-block partial sums of the output block.
struct DiagPartial {
  double energy, ek, ep, px, py, max_u, max_v, max_a, th_max, th_min, ub, vb;
  int64_t max_u_orig, max_v_orig;
};

// Kernel launchers (lps_kernels.hip). fast selects the tile kernels (a.tile must be set).
hipError_t launch_predict(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                          const double *vy, const double *ax, const double *ay, double *px,
                          double *py, hipStream_t s);
// Dilatation pass of the nodes: exact path all nodes and the ABC rows; fast path the tiles (the
// selected kernel on the full tiles, then the unrolled masked kernel on the others).
hipError_t launch_theta(const StepArgs &a, const StepState &st, bool fast, hipStream_t s,
                        bool with_part = true);
// The partial tiles' dilatation kernel when the register-blocked family is selected (else nothing:
// the other families cover every tile in launch_theta).
hipError_t launch_theta_part(const StepArgs &a, const StepState &st, hipStream_t s);
hipError_t launch_idw(const StepArgs &a, const StepState &st, hipStream_t s);
// Force pass: exact path all nodes and ABC rows. Fast path: launch_force = the selected kernel
// (force at every node, ABC nodes included, with the Verlet epilogue), launch_force_part = the
// partial tiles when the selected kernel takes only the full ones.
hipError_t launch_force(const StepArgs &a, const StepState &st, bool fast, hipStream_t s);
// predict_abc: also the ABC nodes' next prediction from U^{n+1}, V^{n+1} (G_u / G_v done) and the new
// acceleration, as k_abc_predict (all ABC nodes must lie in partial tiles).
hipError_t launch_force_part(const StepArgs &a, const StepState &st, hipStream_t s,
                             bool predict_abc = false);
// Fast path, ABC rows. launch_abc_theta (G_theta / direct IDW value) reads only the theta history
// and the prediction, launch_abc_uv (G_u, G_v) only the previous steps' state: both may run on a
// second stream beside the tile passes; the force pass needs abc_theta, launch_abc_predict (the
// next prediction at the ABC nodes) needs abc_uv and the force pass.
// The dilatation with the ABC rows' G_theta and G_u / G_v in the same launch (extra workgroups of
// the register-blocked kernel; other families: the row kernels then the dilatation, on s).
hipError_t launch_theta_rows(const StepArgs &a, const StepState &st, const AbcPack &p, hipStream_t s);
// Register-blocked family on a list of full tiles: the dilatation of the list and of every partial
// tile (launch_theta_list), the force of the list (launch_force_list), and the dilatation + force
// of deep tiles in one launch (launch_force_deep: a full tile whose 3 x 3 neighbourhood is full
// and whose 24 x 24 window holds interior nodes only; it reads u alone, so it needs no dilatation
// pass in front of it, and writes its tile's theta for the other tiles' force).
hipError_t launch_theta_list(const StepArgs &a, const StepState &st, const int32_t *list,
                             int32_t count, hipStream_t s);
hipError_t launch_force_list(const StepArgs &a, const StepState &st, const int32_t *list,
                             int32_t count, hipStream_t s);
hipError_t launch_force_deep(const StepArgs &a, const StepState &st, const int32_t *list,
                             int32_t count, hipStream_t s);
// The n members' packed state (slot: their slots; th null without G_theta): the level of U^n, V^n
// into cur, and with prev not null also U^{n-1}, V^{n-1} into prev (after a reset, a restore, or
// at a step that starts a captured graph); U of an interior member (kind 1) from the prediction
// ring r, of an ABC member from U. Before the step's row kernels.
hipError_t launch_abc_pack(int32_t n, const int32_t *slot, const int8_t *kind, const StepState &st,
                           const RingU &r, double4 *cur, double4 *prev, double2 *th, hipStream_t s);
hipError_t launch_abc_theta(const StepArgs &a, const StepState &st, const AbcPack &p, hipStream_t s);
hipError_t launch_abc_uv(const StepArgs &a, const StepState &st, const AbcPack &p, hipStream_t s);
hipError_t launch_abc_predict(const StepArgs &a, const StepState &st, hipStream_t s);
// U at the interior nodes (kind 1) copied from the prediction ring: ux[i] = px[i] there.
hipError_t launch_u_from_ring(const StepArgs &a, const double *px, const double *py, double *ux,
                              double *uy, hipStream_t s);
hipError_t launch_theta_field(const StepArgs &a, const double *ux, const double *uy, double *theta,
                              hipStream_t s);
hipError_t launch_weighted_volume(const StepArgs &a, double *m, hipStream_t s);
hipError_t launch_diag(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                       const double *vy, const double *ax, const double *ay, const double *theta,
                       const int32_t *perm, double exclude_r, DiagPartial *partials,
                       int32_t n_partials, DiagPartial *out, hipStream_t s);
int32_t diag_partials(int32_t n);
// Fast-path output block: one block per tile (n_partials = a.tile.n_tiles), the strain energy
// from the LDS-staged window and the stencil table.
hipError_t launch_diag_tile(const StepArgs &a, const double *ux, const double *uy,
                            const double *vx, const double *vy, const double *ax, const double *ay,
                            const double *theta, const int32_t *perm, DiagPartial *partials,
                            DiagPartial *out, hipStream_t s);
// Per instance: out[i] reduces the tile partials [tile_off[i], tile_off[i + 1]) in the order of
// launch_diag_tile's final reduction (the same bits for an instance alone or in a batch).
hipError_t launch_diag_final_inst(const DiagPartial *partials, const int32_t *tile_off,
                                  int32_t n_inst, DiagPartial *out, hipStream_t s);
// Tile path check: for every node slot, the stencil cells that hold a node are exactly the set of
// its CSR row (self excluded). bond_of: (2 halo + 1)^2 entries, the bond of offset (ox, oy) at
// (ox + halo) (2 halo + 1) + oy + halo, -1 off the stencil. Adds the number of rows that differ to
// *d_mismatch.
hipError_t launch_tile_check(const StepArgs &a, const int16_t *bond_of, int32_t *d_mismatch,
                             hipStream_t s);

// Horizon build (lps_neighbors.hip) over slots: the CSR and SELL of the nodes' horizons.
struct HorizonBuild {
  int64_t nnz = 0;                 // CSR entries including self
  int64_t *row_ptr = nullptr;      // n_slots + 1
  int32_t *col = nullptr;          // nnz (slots)
  int64_t *sell_off = nullptr;     // slices + 1
  int32_t *sell_w = nullptr;       // slices
  int32_t *sell_col = nullptr;
  int64_t sell_len = 0;
  int32_t max_len = 0;
};
// Morton order of the nodes (non-ABC by position, then seg == 1 nodes by seg_rank): perm[k] is
// the original index of slot k.
hipError_t morton_order(int32_t n, const double *xo, const double *yo, const int8_t *seg,
                        const int32_t *seg_rank, double cell, int32_t *perm, hipStream_t s);
// xs, ys: slot coordinates; kind: slot kind (0 = hole); perm: slot -> original index (ties of
// equal distance are ranked by it). *overflow is set when a horizon exceeds max_len (before the
// notches cut any bond, as in the reference). notch: n_notch x 4 device array of prenotch
// segments (x0, y0, x1, y1); the bonds they cut are left out of the rows.
hipError_t build_horizons(int32_t n_slots, const double *xs, const double *ys, const int8_t *kind,
                          const int32_t *perm, double r, int32_t max_len, const double *notch,
                          int32_t n_notch, HorizonBuild *hb, bool *overflow, hipStream_t s);
void free_horizons(HorizonBuild *hb);

// Message of the last failed call that has no context to hold it (lps_global_error).
extern thread_local std::string g_create_error;

}  // namespace lps
