// Horizon (neighbour list) build on the GPU.
//
// Replaces build_clouds (src/helpers/kd_tree_interface.h:41-126, called at src/main.cpp:352-354).
// The reference inserts every node into a kd-tree, queries the inclusive box [x - r, x + r]^2
// (libraries/libkdtree++/kdtree++/region.hpp:44-53, 86-95), keeps the nodes with
// sqrt(dx^2 + dy^2) <= r (kd_tree_interface.h:90-94) and sorts them by that distance (:107).
// Here:
//   1. the caller places the nodes in slots (lps_api.hip: lattice tiles, or the Morton order built
//      here with a stable radix sort - non-ABC nodes by the Morton code of their position, ABC
//      nodes after them in ABC-row order, equal keys in original order);
//   2. nodes are binned into square cells of side slightly above r, so the 3x3 cells around a node
//      hold every candidate;
//   3. count pass, scan, fill pass: the same inclusive box test and the same FP64 distance
//      (no contraction, correctly rounded sqrt) as the reference, so the sets are identical;
//   4. each row is ranked by (distance, original id) - one wavefront per row, ranks by counting -
//      which puts self first;
//   5. prenotch visibility (src/main.cpp:386-402): bonds whose segment meets a notch segment are
//      dropped from the sorted rows, the rest keep their order (the reference marks them -1 and
//      every later loop skips them, which is the same);
//   6. the CSR is repacked as SELL-64 without self for the step kernels.
#include <hipcub/hipcub.hpp>

#include "lps_internal.h"

namespace lps {

__device__ __forceinline__ uint64_t spread2(uint32_t v) {
  // 2D interleave: bit b of v -> bit 2b
  uint64_t x = v & 0xffffffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

__global__ void k_morton_keys(int32_t n, const double *xo, const double *yo, const int8_t *seg,
                              const int32_t *seg_rank, double xmin, double ymin, double q,
                              uint64_t *keys, int32_t *vals) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t key;
  if (seg && seg[i]) {
    key = (1ull << 62) | (uint64_t)(uint32_t)seg_rank[i];
  } else {
    const double fx = floor((xo[i] - xmin) / q), fy = floor((yo[i] - ymin) / q);
    const uint32_t qx = (uint32_t)fmin(fmax(fx, 0.0), 2097151.0);
    const uint32_t qy = (uint32_t)fmin(fmax(fy, 0.0), 2097151.0);
    key = (spread2(qx) << 1) | spread2(qy);  // 42 bits
  }
  keys[i] = key;
  vals[i] = i;
}

struct CellGrid {
  double xmin, ymin, h;
  int32_t ncx, ncy;
};

__device__ __forceinline__ void cell_of(const CellGrid &g, double x, double y, int32_t *cx,
                                        int32_t *cy) {
  *cx = (int32_t)floor((x - g.xmin) / g.h);
  *cy = (int32_t)floor((y - g.ymin) / g.h);
}

__global__ void k_cell_count(int32_t n, const double *xn, const double *yn, const int8_t *kind,
                             CellGrid g, int32_t *cell_cnt) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || kind[k] == 0) return;
  int32_t cx, cy;
  cell_of(g, xn[k], yn[k], &cx, &cy);
  atomicAdd(&cell_cnt[(int64_t)cx * g.ncy + cy], 1);
}

__global__ void k_cell_fill(int32_t n, const double *xn, const double *yn, const int8_t *kind,
                            CellGrid g, int32_t *cell_fill, int32_t *cell_nodes) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || kind[k] == 0) return;
  int32_t cx, cy;
  cell_of(g, xn[k], yn[k], &cx, &cy);
  const int32_t pos = atomicAdd(&cell_fill[(int64_t)cx * g.ncy + cy], 1);
  cell_nodes[pos] = k;
}

// Inclusive box of the kd-tree query, then the reference's Euclidean test.
__device__ __forceinline__ bool in_horizon(double xi, double yi, double lox, double hix,
                                           double loy, double hiy, double xj, double yj, double r,
                                           double *dist) {
#pragma clang fp contract(off)
  if (xj < lox || hix < xj || yj < loy || hiy < yj) return false;
  const double x = xj - xi;
  const double y = yj - yi;
  const double d = sqrt(x * x + y * y);
  *dist = d;
  return d <= r;
}

template <bool FILL>
__global__ void k_search(int32_t n, const double *xn, const double *yn, const int8_t *kind,
                         CellGrid g, const int32_t *cell_start, const int32_t *cell_nodes, double r,
                         int32_t max_len, int64_t *len_or_ptr, int32_t *cand_col,
                         double *cand_dist, int32_t *overflow) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || kind[k] == 0) return;
  double xi, yi, lox, hix, loy, hiy;
  {
#pragma clang fp contract(off)
    xi = xn[k];
    yi = yn[k];
    lox = xi - r;
    hix = xi + r;
    loy = yi - r;
    hiy = yi + r;
  }
  int32_t cx, cy;
  cell_of(g, xi, yi, &cx, &cy);
  int64_t out = FILL ? len_or_ptr[k] : 0;
  int32_t cnt = 0;
  for (int32_t ax = max(cx - 1, 0); ax <= min(cx + 1, g.ncx - 1); ax++) {
    for (int32_t ay = max(cy - 1, 0); ay <= min(cy + 1, g.ncy - 1); ay++) {
      const int64_t c = (int64_t)ax * g.ncy + ay;
      for (int32_t q = cell_start[c]; q < cell_start[c + 1]; q++) {
        const int32_t j = cell_nodes[q];
        double d;
        if (in_horizon(xi, yi, lox, hix, loy, hiy, xn[j], yn[j], r, &d)) {
          if (FILL) {
            cand_col[out + cnt] = j;
            cand_dist[out + cnt] = d;
          }
          cnt++;
        }
      }
    }
  }
  if (!FILL) {
    len_or_ptr[k] = cnt;
    if (cnt > max_len) atomicExch(overflow, 1);
  }
}

// Side of (px, py) relative to the directed segment (x1, y1) -> (x2, y2): the sign of the cross
// product; on the segment's line -1 before its start, +1 past its end, 0 on it
// (src/main.cpp:119-132, the same FP64 operations without contraction).
__device__ __forceinline__ int notch_side(double x1, double y1, double x2, double y2, double px,
                                          double py) {
#pragma clang fp contract(off)
  const double dx = x2 - x1, dy = y2 - y1;
  double qx = px - x1, qy = py - y1;
  double c = qx * dy - qy * dx;
  if (c == 0.0) {
    c = qx * dx + qy * dy;
    if (c > 0.0) {
      qx -= dx;
      qy -= dy;
      c = qx * dx + qy * dy;
      if (c < 0.0) c = 0.0;
    }
  }
  return (c > 0.0) - (c < 0.0);
}

// The bond i -> j meets (or touches) notch segment k (src/main.cpp:134-140).
__device__ __forceinline__ bool bond_cut(double xi, double yi, double xj, double yj,
                                         const double *nt) {
  return notch_side(xi, yi, xj, yj, nt[0], nt[1]) * notch_side(xi, yi, xj, yj, nt[2], nt[3]) <= 0 &&
         notch_side(nt[0], nt[1], nt[2], nt[3], xi, yi) *
                 notch_side(nt[0], nt[1], nt[2], nt[3], xj, yj) <= 0;
}

// Rows with the notched bonds dropped: FILL = false writes each row's new length, FILL = true
// copies the kept entries in order to their new place. One thread per row; self (entry 0) stays.
template <bool FILL>
__global__ void k_notch_rows(int32_t n, const double *xn, const double *yn, const int64_t *row_ptr,
                             const int32_t *col, const double *notch, int32_t n_notch,
                             int64_t *len_or_ptr, int32_t *col_out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = row_ptr[i], e = row_ptr[i + 1];
  const double xi = xn[i], yi = yn[i];
  int64_t out = FILL ? len_or_ptr[i] : 0;
  int32_t kept = 0;
  for (int64_t q = b; q < e; q++) {
    const int32_t j = col[q];
    bool cut = false;
    if (q > b)
      for (int32_t k = 0; k < n_notch && !cut; k++) cut = bond_cut(xi, yi, xn[j], yn[j], notch + 4 * k);
    if (cut) continue;
    if (FILL) col_out[out + kept] = j;
    kept++;
  }
  if (!FILL) len_or_ptr[i] = kept;
}

// Rank each row by (distance, original id): one wavefront per row, ranks by counting.
__global__ void __launch_bounds__(256) k_sort_rows(int32_t n, const int64_t *row_ptr,
                                                   const int32_t *cand_col,
                                                   const double *cand_dist, const int32_t *perm,
                                                   int32_t *col) {
  __shared__ double sd[4][kMaxHorizon];
  __shared__ int32_t so[4][kMaxHorizon];
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int32_t row = blockIdx.x * 4 + wid;
  if (row >= n) return;
  const int64_t b = row_ptr[row];
  const int32_t len = (int32_t)(row_ptr[row + 1] - b);
  for (int t = lane; t < len; t += 64) {
    sd[wid][t] = cand_dist[b + t];
    so[wid][t] = perm[cand_col[b + t]];
  }
  __threadfence_block();
  for (int t = lane; t < len; t += 64) {
    const double d = sd[wid][t];
    const int32_t o = so[wid][t];
    int32_t rank = 0;
    for (int f = 0; f < len; f++) {
      const double df = sd[wid][f];
      rank += (df < d) || (df == d && so[wid][f] < o);
    }
    col[b + rank] = cand_col[b + t];
  }
}

__global__ void k_slice_width(int32_t n, int32_t nslices, const int64_t *row_ptr, int32_t *w,
                              int64_t *wsz) {
  const int32_t lane = threadIdx.x & 63;
  const int32_t s = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int32_t row = s * kSlice + lane;
  int32_t len = 0;
  if (row < n) len = max((int32_t)(row_ptr[row + 1] - row_ptr[row]) - 1, 0);  // self excluded
  for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, 64));
  if (lane == 0) {
    w[s] = len;
    wsz[s] = (int64_t)len * kSlice;
  }
}

__global__ void k_sell_fill(int32_t n, const int64_t *row_ptr, const int32_t *col,
                            const int64_t *sell_off, const int32_t *sell_w, int32_t *sell_col) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t s = k / kSlice, lane = k % kSlice;
  const int64_t b = row_ptr[k] + 1, len = row_ptr[k + 1] - b;  // holes: len = -1, all pads
  int32_t *out = sell_col + sell_off[s] + lane;
  const int32_t w = sell_w[s];
  for (int32_t t = 0; t < w; t++) out[(int64_t)t * kSlice] = (t < len) ? col[b + t] : -1;
  (void)n;
}

__global__ void k_max_len(int32_t n, const int64_t *row_ptr, int32_t *out) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  atomicMax(out, (int32_t)(row_ptr[k + 1] - row_ptr[k]));
}

#define LPS_TRY(x)                           \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) return e_;         \
  } while (0)

static inline unsigned grid_of(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

void free_horizons(HorizonBuild *hb) {
  (void)hipFree(hb->row_ptr);
  (void)hipFree(hb->col); (void)hipFree(hb->sell_off); (void)hipFree(hb->sell_w);
  (void)hipFree(hb->sell_col);
  *hb = HorizonBuild();
}

namespace {
// Device buffers released when the build returns.
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t get(T **p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};
}  // namespace

hipError_t morton_order(int32_t n, const double *xo, const double *yo, const int8_t *seg,
                        const int32_t *seg_rank, double cell, int32_t *perm, hipStream_t s) {
  Scratch sc;
  std::vector<double> hx(n), hy(n);
  LPS_TRY(hipMemcpyAsync(hx.data(), xo, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  LPS_TRY(hipMemcpyAsync(hy.data(), yo, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  LPS_TRY(hipStreamSynchronize(s));
  double xmin = hx[0], ymin = hy[0];
  for (int32_t i = 1; i < n; i++) {
    xmin = std::min(xmin, hx[i]);
    ymin = std::min(ymin, hy[i]);
  }
  uint64_t *keys, *keys_out;
  int32_t *vals;
  LPS_TRY(sc.get(&keys, n));
  LPS_TRY(sc.get(&keys_out, n));
  LPS_TRY(sc.get(&vals, n));
  hipLaunchKernelGGL(k_morton_keys, dim3(grid_of(n)), dim3(256), 0, s, n, xo, yo, seg, seg_rank,
                     xmin, ymin, cell / 8.0, keys, vals);
  LPS_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  LPS_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_out, vals, perm, n, 0,
                                             64, s));
  char *tmp;
  LPS_TRY(sc.get(&tmp, tmp_bytes));
  LPS_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_out, vals, perm, n, 0, 64,
                                             s));
  LPS_TRY(hipStreamSynchronize(s));
  return hipSuccess;
}

hipError_t build_horizons(int32_t n, const double *xn, const double *yn, const int8_t *kind,
                          const int32_t *perm, double r, int32_t max_len, const double *notch,
                          int32_t n_notch, HorizonBuild *hb, bool *overflow, hipStream_t s) {
  *overflow = false;
  Scratch sc;
  // bounding box of the nodes (holes excluded)
  std::vector<double> hx(n), hy(n);
  std::vector<int8_t> hk(n);
  LPS_TRY(hipMemcpyAsync(hx.data(), xn, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  LPS_TRY(hipMemcpyAsync(hy.data(), yn, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  LPS_TRY(hipMemcpyAsync(hk.data(), kind, n, hipMemcpyDeviceToHost, s));
  LPS_TRY(hipStreamSynchronize(s));
  double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
  bool first = true;
  for (int32_t i = 0; i < n; i++) {
    if (!hk[i]) continue;
    if (first) {
      xmin = xmax = hx[i];
      ymin = ymax = hy[i];
      first = false;
    }
    xmin = std::min(xmin, hx[i]); xmax = std::max(xmax, hx[i]);
    ymin = std::min(ymin, hy[i]); ymax = std::max(ymax, hy[i]);
  }
  CellGrid g;
  g.xmin = xmin;
  g.ymin = ymin;
  g.h = r * (1.0 + 1e-7);  // margin against rounding in the cell index
  g.ncx = (int32_t)std::floor((xmax - xmin) / g.h) + 1;
  g.ncy = (int32_t)std::floor((ymax - ymin) / g.h) + 1;
  const int64_t ncell = (int64_t)g.ncx * g.ncy;
  if (ncell > (int64_t)1 << 30) return hipErrorInvalidValue;

  // cell binning
  int32_t *cell_cnt, *cell_start, *cell_fill, *cell_nodes;
  LPS_TRY(sc.get(&cell_cnt, ncell + 1));
  LPS_TRY(sc.get(&cell_start, ncell + 1));
  LPS_TRY(sc.get(&cell_fill, ncell + 1));
  LPS_TRY(sc.get(&cell_nodes, n));
  LPS_TRY(hipMemsetAsync(cell_cnt, 0, sizeof(int32_t) * (ncell + 1), s));
  hipLaunchKernelGGL(k_cell_count, dim3(grid_of(n)), dim3(256), 0, s, n, xn, yn, kind, g, cell_cnt);
  LPS_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  LPS_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cell_cnt, cell_start, ncell + 1, s));
  char *tmp2;
  LPS_TRY(sc.get(&tmp2, tmp_bytes));
  LPS_TRY(hipcub::DeviceScan::ExclusiveSum(tmp2, tmp_bytes, cell_cnt, cell_start, ncell + 1, s));
  LPS_TRY(hipMemcpyAsync(cell_fill, cell_start, sizeof(int32_t) * (ncell + 1),
                         hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_cell_fill, dim3(grid_of(n)), dim3(256), 0, s, n, xn, yn, kind, g, cell_fill,
                     cell_nodes);
  LPS_TRY(hipGetLastError());

  // count, scan, fill
  int64_t *lens;
  int32_t *d_over;
  LPS_TRY(sc.get(&lens, n + 1));
  LPS_TRY(sc.get(&d_over, 1));
  LPS_TRY(hipMemsetAsync(d_over, 0, sizeof(int32_t), s));
  LPS_TRY(hipMemsetAsync(lens, 0, sizeof(int64_t) * (n + 1), s));
  hipLaunchKernelGGL(k_search<false>, dim3(grid_of(n)), dim3(256), 0, s, n, xn, yn, kind, g,
                     cell_start, cell_nodes, r, max_len, lens, nullptr, nullptr, d_over);
  LPS_TRY(hipGetLastError());
  int32_t h_over = 0;
  LPS_TRY(hipMemcpyAsync(&h_over, d_over, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  LPS_TRY(hipStreamSynchronize(s));
  if (h_over) {
    *overflow = true;
    return hipSuccess;
  }
  LPS_TRY(hipMalloc((void **)&hb->row_ptr, sizeof(int64_t) * (n + 1)));
  tmp_bytes = 0;
  LPS_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, lens, hb->row_ptr, n + 1, s));
  char *tmp3;
  LPS_TRY(sc.get(&tmp3, tmp_bytes));
  LPS_TRY(hipcub::DeviceScan::ExclusiveSum(tmp3, tmp_bytes, lens, hb->row_ptr, n + 1, s));
  LPS_TRY(hipMemcpyAsync(&hb->nnz, hb->row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  LPS_TRY(hipStreamSynchronize(s));
  int32_t *cand_col;
  double *cand_dist;
  LPS_TRY(sc.get(&cand_col, hb->nnz));
  LPS_TRY(sc.get(&cand_dist, hb->nnz));
  hipLaunchKernelGGL(k_search<true>, dim3(grid_of(n)), dim3(256), 0, s, n, xn, yn, kind, g,
                     cell_start, cell_nodes, r, max_len, hb->row_ptr, cand_col, cand_dist, d_over);
  LPS_TRY(hipGetLastError());

  // rank rows by (distance, original index)
  LPS_TRY(hipMalloc((void **)&hb->col, sizeof(int32_t) * std::max<int64_t>(hb->nnz, 1)));
  hipLaunchKernelGGL(k_sort_rows, dim3(grid_of(n, 4)), dim3(256), 0, s, n, hb->row_ptr, cand_col,
                     cand_dist, perm, hb->col);
  LPS_TRY(hipGetLastError());

  // prenotch visibility: count the kept bonds per row, scan, copy them in order
  if (n_notch > 0) {
    int64_t *klen, *kptr;
    int32_t *kcol;
    LPS_TRY(sc.get(&klen, n + 1));
    LPS_TRY(hipMemsetAsync(klen, 0, sizeof(int64_t) * (n + 1), s));
    hipLaunchKernelGGL(k_notch_rows<false>, dim3(grid_of(n)), dim3(256), 0, s, n, xn, yn,
                       hb->row_ptr, hb->col, notch, n_notch, klen, nullptr);
    LPS_TRY(hipGetLastError());
    LPS_TRY(hipMalloc((void **)&kptr, sizeof(int64_t) * (n + 1)));
    tmp_bytes = 0;
    LPS_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, klen, kptr, n + 1, s));
    char *tmp5;
    LPS_TRY(sc.get(&tmp5, tmp_bytes));
    LPS_TRY(hipcub::DeviceScan::ExclusiveSum(tmp5, tmp_bytes, klen, kptr, n + 1, s));
    int64_t knnz = 0;
    LPS_TRY(hipMemcpyAsync(&knnz, kptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LPS_TRY(hipStreamSynchronize(s));
    LPS_TRY(hipMalloc((void **)&kcol, sizeof(int32_t) * std::max<int64_t>(knnz, 1)));
    hipLaunchKernelGGL(k_notch_rows<true>, dim3(grid_of(n)), dim3(256), 0, s, n, xn, yn,
                       hb->row_ptr, hb->col, notch, n_notch, kptr, kcol);
    LPS_TRY(hipGetLastError());
    LPS_TRY(hipStreamSynchronize(s));
    (void)hipFree(hb->row_ptr);
    (void)hipFree(hb->col);
    hb->row_ptr = kptr;
    hb->col = kcol;
    hb->nnz = knnz;
  }

  // SELL-64 without self
  const int32_t nslices = (n + kSlice - 1) / kSlice;
  int64_t *wsz;
  LPS_TRY(sc.get(&wsz, nslices + 1));
  LPS_TRY(hipMalloc((void **)&hb->sell_w, sizeof(int32_t) * (nslices + 1)));
  LPS_TRY(hipMalloc((void **)&hb->sell_off, sizeof(int64_t) * (nslices + 1)));
  LPS_TRY(hipMemsetAsync(wsz, 0, sizeof(int64_t) * (nslices + 1), s));
  hipLaunchKernelGGL(k_slice_width, dim3(grid_of(nslices, 4)), dim3(256), 0, s, n, nslices,
                     hb->row_ptr, hb->sell_w, wsz);
  LPS_TRY(hipGetLastError());
  tmp_bytes = 0;
  LPS_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, wsz, hb->sell_off, nslices + 1, s));
  char *tmp4;
  LPS_TRY(sc.get(&tmp4, tmp_bytes));
  LPS_TRY(hipcub::DeviceScan::ExclusiveSum(tmp4, tmp_bytes, wsz, hb->sell_off, nslices + 1, s));
  LPS_TRY(hipMemcpyAsync(&hb->sell_len, hb->sell_off + nslices, sizeof(int64_t),
                         hipMemcpyDeviceToHost, s));
  LPS_TRY(hipStreamSynchronize(s));
  LPS_TRY(hipMalloc((void **)&hb->sell_col, sizeof(int32_t) * std::max<int64_t>(hb->sell_len, 1)));
  LPS_TRY(hipMemsetAsync(hb->sell_col, 0xff, sizeof(int32_t) * std::max<int64_t>(hb->sell_len, 1), s));
  hipLaunchKernelGGL(k_sell_fill, dim3(grid_of(n)), dim3(256), 0, s, n, hb->row_ptr, hb->col,
                     hb->sell_off, hb->sell_w, hb->sell_col);
  LPS_TRY(hipGetLastError());
  int32_t *d_max;
  LPS_TRY(sc.get(&d_max, 1));
  LPS_TRY(hipMemsetAsync(d_max, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_max_len, dim3(grid_of(n)), dim3(256), 0, s, n, hb->row_ptr, d_max);
  LPS_TRY(hipGetLastError());
  LPS_TRY(hipMemcpyAsync(&hb->max_len, d_max, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  LPS_TRY(hipStreamSynchronize(s));
  return hipSuccess;
}

}  // namespace lps
