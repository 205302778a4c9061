// field.hip -- per-tessellation setup on the GPU (gfx950):
//   validation of the uploaded tets, tet volumes, vertex -> tet CSR, vertex
//   density rho_i = 4 m / W_i, per-tet gradients, packed walk records, and
//   the forward-hull entry grid.
//
// Arithmetic that decides results is written in the reference's operation
// order and built with -ffp-contract=off, so volumes, densities and
// gradients equal the reference's bit for bit:
//   volume    src/triangulate.c:190-215
//   density   src/triangulate.c:217-224 (W summed in ascending tet order)
//   gradient  src/triangulate.c:228-257
//   forward   src/triangulate.c:274-311
//   triangles src/triangulate.c:325-355 (order: tet, then face BDC..ABC)
#include <algorithm>
#include <cmath>
#include <vector>

#include "sdtfe_host.h"
#include "walk_common.h"
#include "walk_fp64_face.h"

namespace sdtfe {

// ============================================================ scan

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ inline int wave_inclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int n = __shfl_up(v, off, 64);
    if (lane >= off) v += n;
  }
  return v;
}

// Exclusive scan over the block; wsum holds blockDim/64 + 1 ints.
__device__ inline int block_exclusive(int v, int *wsum, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int inc = wave_inclusive(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < nw; ++i) {
      int s = wsum[i];
      wsum[i] = acc;
      acc += s;
    }
    wsum[nw] = acc;
  }
  __syncthreads();
  int ex = inc - v + wsum[w];
  *total = wsum[nw];
  __syncthreads();
  return ex;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const int *in, int64_t n,
                                                                    int *partial) {
  __shared__ int wsum[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  int tot;
  block_exclusive(s, wsum, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_partials_kernel(int *partial, int nb, int *total) {
  __shared__ int wsum[1024 / 64 + 1];
  const int per = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
  const int b0 = threadIdx.x * per;
  int s = 0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < nb) s += partial[b0 + i];
  int tot;
  int ex = block_exclusive(s, wsum, &tot);
  for (int i = 0; i < per; ++i)
    if (b0 + i < nb) {
      int v = partial[b0 + i];
      partial[b0 + i] = ex;
      ex += v;
    }
  if (threadIdx.x == 0) *total = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const int *in, int *out,
                                                                   int64_t n,
                                                                   const int *partial) {
  __shared__ int wsum[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  int tot;
  int ex = block_exclusive(s, wsum, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
}

size_t scan_tmp_ints(int64_t n) { return (size_t)((n + kScanChunk - 1) / kScanChunk) + 1; }

// Where scan_exclusive leaves the total of n > 0 items in its tmp space.
const int *scan_total_device(const int *tmp, int64_t n) {
  return tmp + (n + kScanChunk - 1) / kScanChunk;
}

int scan_exclusive(const int *in, int *out, int64_t n, int *tmp, int *total_host,
                   hipStream_t s) {
  if (n <= 0) {
    if (total_host) *total_host = 0;
    return 0;
  }
  const int64_t nb = (n + kScanChunk - 1) / kScanChunk;
  if (nb > 1024 * 64) return fail(SDTFE_EINVAL, "scan: too many items");
  int *partial = tmp, *total = tmp + nb;
  scan_reduce_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, partial);
  scan_partials_kernel<<<1, 1024, 0, s>>>(partial, (int)nb, total);
  scan_apply_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, partial);
  SDTFE_HIP(hipGetLastError());
  if (total_host) {
    SDTFE_HIP(hipMemcpyAsync(total_host, total, sizeof(int), hipMemcpyDeviceToHost, s));
    SDTFE_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

// ============================================================ field

__device__ inline int4 tet_verts(const int *tets, int t) {
  return reinterpret_cast<const int4 *>(tets)[2 * (int64_t)t];
}
__device__ inline int4 tet_nbrs(const int *tets, int t) {
  return reinterpret_cast<const int4 *>(tets)[2 * (int64_t)t + 1];
}

__global__ void validate_kernel(const int *tets, int nt, int n, int *bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int4 v = tet_verts(tets, t), nb = tet_nbrs(tets, t);
  bool ok = v.x >= 0 && v.x < n && v.y >= 0 && v.y < n && v.z >= 0 && v.z < n && v.w >= 0 &&
            v.w < n && nb.x >= -1 && nb.x < nt && nb.y >= -1 && nb.y < nt && nb.z >= -1 &&
            nb.z < nt && nb.w >= -1 && nb.w < nt;
  if (!ok) atomicOr(bad, 1);
}

// (A - D) . ((B - D) x (C - D)) in the reference's expansion order.
__device__ inline double vol6(const double4 &a, const double4 &b, const double4 &c,
                              const double4 &d) {
  double m0 = (b.y - d.y) * (c.z - d.z) - (b.z - d.z) * (c.y - d.y);
  double m1 = (b.z - d.z) * (c.x - d.x) - (b.x - d.x) * (c.z - d.z);
  double m2 = (b.x - d.x) * (c.y - d.y) - (b.y - d.y) * (c.x - d.x);
  return (a.x - d.x) * m0 + (a.y - d.y) * m1 + (a.z - d.z) * m2;
}

__global__ void volume_degree_kernel(const double4 *pos, const int *tets, int nt, double *vol,
                                     int *deg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int4 v = tet_verts(tets, t);
  vol[t] = fabs(vol6(pos[v.x], pos[v.y], pos[v.z], pos[v.w]) / 6);
  atomicAdd(deg + v.x, 1);
  atomicAdd(deg + v.y, 1);
  atomicAdd(deg + v.z, 1);
  atomicAdd(deg + v.w, 1);
}

__global__ void fill_incidence_kernel(const int *tets, int nt, const int *off, int *cursor,
                                      int *inc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int4 v = tet_verts(tets, t);
  inc[off[v.x] + atomicAdd(cursor + v.x, 1)] = t;
  inc[off[v.y] + atomicAdd(cursor + v.y, 1)] = t;
  inc[off[v.z] + atomicAdd(cursor + v.z, 1)] = t;
  inc[off[v.w] + atomicAdd(cursor + v.w, 1)] = t;
}

// One wave per vertex: the volumes of its incident tets added in ascending
// tet order, the order the reference adds them (it loops over the tets).
// Up to 64 incident tets: lane k holds the k-th, finds its rank among them
// across the wave (ties, from a tet naming a vertex twice, by lane), and the
// volumes are added rank by rank.  Beyond 64 (rare), one lane puts them in
// order in place and adds them.
__global__ __launch_bounds__(256) void vertex_density_kernel(double4 *pos, int n, const int *off,
                                                             int *inc, const double *vol,
                                                             float m4) {
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (v >= n) return;  // the same for the whole wave
  const int b = off[v], e = off[v + 1], d = e - b;
  double w = 0;
  if (d <= 64) {
    const int key = lane < d ? inc[b + lane] : INT_MAX;
    int rank = 0;
    for (int k = 0; k < d; ++k) {
      const int o = __shfl(key, k, 64);
      rank += o < key || (o == key && k < lane);
    }
    const double vv = lane < d ? vol[key] : 0.0;
    for (int r = 0; r < d; ++r) {
      const unsigned long long m = __ballot(lane < d && rank == r);
      w += __shfl(vv, __ffsll(m) - 1, 64);
    }
  } else if (lane == 0) {
    for (int i = b + 1; i < e; ++i) {
      const int key = inc[i];
      int j = i - 1;
      while (j >= b && inc[j] > key) {
        inc[j + 1] = inc[j];
        --j;
      }
      inc[j + 1] = key;
    }
    for (int i = b; i < e; ++i) w += vol[inc[i]];
  }
  if (lane == 0) pos[v].w = (double)m4 / w;
}

__device__ inline double det3(double a, double b, double c, double d, double e, double f,
                              double g, double h, double i) {
  return a * e * i + b * f * g + c * d * h - c * e * g - b * d * i - a * f * h;
}

__global__ void gradient_record_kernel(const double4 *pos, const int *tets, int nt,
                                       double *nabla, Rec64 *rec, Rec32 *rec32) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int4 v = tet_verts(tets, t), nb = tet_nbrs(tets, t);
  const double4 a = pos[v.x], b = pos[v.y], c = pos[v.z], d = pos[v.w];
  const double xb = b.x - a.x, yb = b.y - a.y, zb = b.z - a.z;
  const double xc = c.x - a.x, yc = c.y - a.y, zc = c.z - a.z;
  const double xd = d.x - a.x, yd = d.y - a.y, zd = d.z - a.z;
  const double rb = b.w - a.w, rc = c.w - a.w, rd = d.w - a.w;
  const double den = det3(xb, yb, zb, xc, yc, zc, xd, yd, zd);
  const double gx = det3(rb, yb, zb, rc, yc, zc, rd, yd, zd) / den;
  const double gy = det3(xb, rb, zb, xc, rc, zc, xd, rd, zd) / den;
  const double gz = det3(xb, yb, rb, xc, yc, rc, xd, yd, rd) / den;
  nabla[3 * (int64_t)t + 0] = gx;
  nabla[3 * (int64_t)t + 1] = gy;
  nabla[3 * (int64_t)t + 2] = gz;

  int4 *ri = reinterpret_cast<int4 *>(rec + t);
  ri[0] = nb;
  double *rd64 = reinterpret_cast<double *>(ri + 1);
  rd64[0] = a.x; rd64[1] = a.y; rd64[2] = a.z;
  rd64[3] = b.x; rd64[4] = b.y; rd64[5] = b.z;
  rd64[6] = c.x; rd64[7] = c.y; rd64[8] = c.z;
  rd64[9] = d.x; rd64[10] = d.y; rd64[11] = d.z;
  rd64[12] = a.w; rd64[13] = gx; rd64[14] = gy; rd64[15] = gz;

  Rec32 q;
  q.nb[0] = nb.x; q.nb[1] = nb.y; q.nb[2] = nb.z; q.nb[3] = nb.w;
  q.a[0] = a.x; q.a[1] = a.y; q.a[2] = a.z;
  q.e[0] = (float)(b.x - a.x); q.e[1] = (float)(b.y - a.y); q.e[2] = (float)(b.z - a.z);
  q.e[3] = (float)(c.x - a.x); q.e[4] = (float)(c.y - a.y); q.e[5] = (float)(c.z - a.z);
  q.e[6] = (float)(d.x - a.x); q.e[7] = (float)(d.y - a.y); q.e[8] = (float)(d.z - a.z);
  q.rho[0] = (float)a.w; q.rho[1] = (float)b.w; q.rho[2] = (float)c.w; q.rho[3] = (float)d.w;
  q.pad = 0.f;
  rec32[t] = q;
}

// Face records of the mixed walk (sdtfe::FaceRec), thread per (tet, face).
__global__ void face_record_kernel(const double4 *pos, const int *tets, int nt, FaceRec *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * (int64_t)nt) return;
  const int t = (int)(i >> 2), f = (int)(i & 3);
  FaceRec r;
  r.nb = tets[8 * (int64_t)t + 4 + f];
  r.code = kFaceStop;
  r.qxh = r.qxl = r.qyh = r.qyl = 0.f;
  r.dqz = r.rho = 0.f;
  if (r.nb >= 0) {
    const int n = r.nb;
    int back = -1;
    for (int k = 3; k >= 0; --k)
      if (tets[8 * (int64_t)n + 4 + k] == t) back = k;
    // byte f: 8 x back; byte m (m != f): 8 x the label in n of this tet's
    // vertex m
    unsigned code = back < 0 ? kFaceStop : (unsigned)back << (8 * f + 3);
    double zs = 0.0;
    for (int k = 1; k <= 3; ++k) {
      const int v = tets[8 * (int64_t)t + (f ^ k)];
      zs += pos[v].z;
      int lab = -1;
      for (int m = 3; m >= 0; --m)
        if (tets[8 * (int64_t)n + m] == v) lab = m;
      if (lab < 0 || lab == back) code |= kFaceStop;
      code |= (unsigned)(lab & 3) << (8 * (f ^ k) + 3);
    }
    r.code = code;
    const double4 q = pos[tets[8 * (int64_t)n + (back & 3)]];
    split32(q.x, r.qxh, r.qxl);
    split32(q.y, r.qyh, r.qyl);
    r.dqz = (float)(q.z - zs / 3.0);
    r.rho = (float)q.w;
  }
  out[i] = r;
}

// FP64 walk records of every (tet, face).
__global__ void face_record64_kernel(const double4 *pos, const int *tets, const double *nabla,
                                     int nt, FRec64 *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * (int64_t)nt) return;
  FRec64 r;
  make_frec64(pos, tets, nabla, (int)(i >> 2), (int)(i & 3), r);
  out[i] = r;
}

// ============================================================ entry grid

// Face f of a tet as the triangle the reference emits (src/triangulate.c:
// 328-343); the forward test uses (p1 - p0) x (p2 - p0) (:279-308).
__device__ inline void face_triple(int f, int4 v, int &p0, int &p1, int &p2) {
  switch (f) {
    case 0: p0 = v.y; p1 = v.w; p2 = v.z; break;  // BDC
    case 1: p0 = v.x; p1 = v.z; p2 = v.w; break;  // ACD
    case 2: p0 = v.x; p1 = v.w; p2 = v.y; break;  // ADB
    default: p0 = v.x; p1 = v.y; p2 = v.z; break; // ABC
  }
}

// Forward-hull faces of each tet (bit f: face f has no neighbour and faces
// down) and their count.  With part, also each block's bounding box of those
// faces' vertices (x min, y min, x max, y max at part + 4 x block).  Blocks
// of 256 threads.
__global__ __launch_bounds__(256) void hull_mask_kernel(const double4 *pos, const int *tets,
                                                        int nt, unsigned char *mask, int *cnt,
                                                        double *part) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double lo_x = INFINITY, lo_y = INFINITY, hi_x = -INFINITY, hi_y = -INFINITY;
  if (t < nt) {
    const int4 v = tet_verts(tets, t), nb = tet_nbrs(tets, t);
    const int nbf[4] = {nb.x, nb.y, nb.z, nb.w};
    unsigned m = 0;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (nbf[f] != -1) continue;
      int i0, i1, i2;
      face_triple(f, v, i0, i1, i2);
      const double4 p0 = pos[i0], p1 = pos[i1], p2 = pos[i2];
      double t1x = p1.x - p0.x, t1y = p1.y - p0.y;
      double t2x = p2.x - p0.x, t2y = p2.y - p0.y;
      if (t1x * t2y - t1y * t2x < 0.0) {
        m |= 1u << f;
        lo_x = fmin(lo_x, fmin(fmin(p0.x, p1.x), p2.x));
        lo_y = fmin(lo_y, fmin(fmin(p0.y, p1.y), p2.y));
        hi_x = fmax(hi_x, fmax(fmax(p0.x, p1.x), p2.x));
        hi_y = fmax(hi_y, fmax(fmax(p0.y, p1.y), p2.y));
      }
    }
    mask[t] = (unsigned char)m;
    cnt[t] = __popc(m);
  }
  if (!part) return;  // the same for every thread of the launch
  __shared__ double s[4][256];
  s[0][threadIdx.x] = lo_x; s[1][threadIdx.x] = lo_y;
  s[2][threadIdx.x] = hi_x; s[3][threadIdx.x] = hi_y;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + h]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + h]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + h]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int j = 0; j < 4; ++j) part[4 * (int64_t)blockIdx.x + j] = s[j][0];
}

__global__ void emit_tri_kernel(const double4 *pos, const int *tets, int nt,
                                const unsigned char *mask, const int *off, double *tri_xy,
                                int *tri_tet, int *tri_rid, int *tri_abc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const unsigned m = mask[t];
  if (!m) return;
  const int4 v = tet_verts(tets, t);
  int k = off[t];
  for (int f = 0; f < 4; ++f) {
    if (!(m >> f & 1u)) continue;
    int i0, i1, i2;
    face_triple(f, v, i0, i1, i2);
    const double4 p0 = pos[i0], p1 = pos[i1], p2 = pos[i2];
    double *xy = tri_xy + 6 * (int64_t)k;
    xy[0] = p0.x; xy[1] = p0.y; xy[2] = p1.x; xy[3] = p1.y; xy[4] = p2.x; xy[5] = p2.y;
    tri_abc[3 * k + 0] = i0;
    tri_abc[3 * k + 1] = i1;
    tri_abc[3 * k + 2] = i2;
    tri_tet[k] = t;
    tri_rid[k] = nt < (1 << 29) ? 4 * t + f : -1;
    ++k;
  }
}

// The forward hull's bounding box from hull_mask_kernel's nb block boxes,
// and its triangle count (the scan's total): out = {x min, y min, x max,
// y max, count}, one read-back for both.  One block of 256 threads.
__global__ __launch_bounds__(256) void hull_box_kernel(const double *part, int nb,
                                                       const int *count, double *out) {
  __shared__ double s[4][256];
  double lo_x = INFINITY, lo_y = INFINITY, hi_x = -INFINITY, hi_y = -INFINITY;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    const double *b = part + 4 * (int64_t)k;
    lo_x = fmin(lo_x, b[0]);
    lo_y = fmin(lo_y, b[1]);
    hi_x = fmax(hi_x, b[2]);
    hi_y = fmax(hi_y, b[3]);
  }
  s[0][threadIdx.x] = lo_x; s[1][threadIdx.x] = lo_y;
  s[2][threadIdx.x] = hi_x; s[3][threadIdx.x] = hi_y;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + h]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + h]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + h]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int j = 0; j < 4; ++j) out[j] = s[j][0];
    out[4] = (double)*count;
  }
}

template <bool kFill>
__global__ void cell_bin_kernel(EntryGrid g, int *cnt_or_cursor, const int *off, int *cell_tri) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.n_tri) return;
  const double *xy = g.tri_xy + 6 * (int64_t)k;
  double lx = fmin(fmin(xy[0], xy[2]), xy[4]), hx = fmax(fmax(xy[0], xy[2]), xy[4]);
  double ly = fmin(fmin(xy[1], xy[3]), xy[5]), hy = fmax(fmax(xy[1], xy[3]), xy[5]);
  const int cx0 = grid_cell(lx, g.x0, g.sx, g.nx), cx1 = grid_cell(hx, g.x0, g.sx, g.nx);
  const int cy0 = grid_cell(ly, g.y0, g.sy, g.ny), cy1 = grid_cell(hy, g.y0, g.sy, g.ny);
  for (int cy = cy0; cy <= cy1; ++cy)
    for (int cx = cx0; cx <= cx1; ++cx) {
      const int c = cy * g.nx + cx;
      if (kFill)
        cell_tri[off[c] + atomicAdd(cnt_or_cursor + c, 1)] = k;
      else
        atomicAdd(cnt_or_cursor + c, 1);
    }
}

// ============================================================ host side

namespace {

int dev_block(sdtfe_tess *t, void **p, size_t bytes) {
  hipError_t e = dev_malloc(p, bytes, t->stream);
  if (e != hipSuccess) return fail(SDTFE_ENOMEM, "device allocation failed (" +
                                                     std::to_string(bytes) +
                                                     " B): " + hipGetErrorString(e));
  t->device_bytes += (int64_t)bytes;
  return 0;
}

template <typename T>
int dev_alloc(sdtfe_tess *t, T **p, size_t count) {
  return dev_block(t, (void **)p, sizeof(T) * (count > 0 ? count : 1));
}

// Temporaries of one set-up step, freed in stream order when it ends.  A
// step that reserves the sum of its arrays' sizes (size()) first gets them
// carved from one allocation: one allocation and one free instead of one per
// array.  An array that does not fit what is left gets a block of its own.
struct Scratch {
  hipStream_t s;
  std::vector<void *> ptrs;
  char *cur = nullptr;
  size_t left = 0;
  explicit Scratch(hipStream_t stream) : s(stream) {}
  ~Scratch() {
    for (void *p : ptrs) dev_free(p, s);
  }
  template <typename T>
  static size_t size(size_t count) {
    return (sizeof(T) * (count > 0 ? count : 1) + 255) & ~(size_t)255;
  }
  int reserve(size_t bytes) {
    void *p = nullptr;
    if (dev_malloc(&p, bytes, s) != hipSuccess)
      return fail(SDTFE_ENOMEM, "device allocation (scratch) failed");
    ptrs.push_back(p);
    cur = static_cast<char *>(p);
    left = bytes;
    return 0;
  }
  template <typename T>
  int get(T **p, size_t count) {
    const size_t bytes = size<T>(count);
    if (bytes <= left) {
      *p = reinterpret_cast<T *>(cur);
      cur += bytes;
      left -= bytes;
      return 0;
    }
    hipError_t e = dev_malloc((void **)p, bytes, s);
    if (e != hipSuccess) return fail(SDTFE_ENOMEM, "device allocation (scratch) failed");
    ptrs.push_back(*p);
    return 0;
  }
};

// Destination of a small read-back: the first pinned buffer of st when it
// has one (a pageable destination costs the runtime a staged copy of its
// own), else local.  st's uploads were queued on the same stream before, so
// the read-back lands after they have read the buffer.
template <typename T>
T *readback_words(Staging *st, T *local) {
  return st && st->buf[0] ? static_cast<T *>(st->buf[0]) : local;
}

inline unsigned blocks_for(int64_t n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace

#define SDTFE_RC(expr)     \
  do {                     \
    int _rc = (expr);      \
    if (_rc) return _rc;   \
  } while (0)

namespace {

// Consecutive arrays of one allocation, each on a 256-byte boundary; with a
// null base it only adds up their sizes.
struct Carve {
  char *base = nullptr;
  size_t off = 0;
  template <typename T>
  void take(T **p, size_t count) {
    if (base) *p = reinterpret_cast<T *>(base + off);
    off += Scratch::size<T>(count);
  }
};

}  // namespace

int build_field(sdtfe_tess *t, const double *pd, const int *tetra, Staging *st) {
  const int n = t->n, nt = t->nt;
  hipStream_t s = t->stream;
  // the per-particle and per-tet arrays: one allocation, freed with the
  // tessellation (the FP64 walk's face records are addressed as 4 x tet +
  // face in an int, so they are built below 2^29 tets)
  const bool frec64 = t->fp64_faces && nt > 0 && nt < (1 << 29);
  auto layout = [&](Carve &c) {
    c.take(&t->pos4, 4 * (size_t)n);
    c.take(&t->tets, 8 * (size_t)nt);
    c.take(&t->vol, (size_t)nt);
    c.take(&t->nabla, 3 * (size_t)nt);
    c.take(&t->rec64, (size_t)nt);
    c.take(&t->rec32, (size_t)nt);
    c.take(&t->facerec, 4 * (size_t)nt);
    if (frec64) c.take(&t->frec64, 4 * (size_t)nt);
  };
  Carve sizing;
  layout(sizing);
  SDTFE_RC(dev_block(t, &t->field_block, sizing.off));
  Carve carve{static_cast<char *>(t->field_block)};
  layout(carve);
  SDTFE_RC(upload(t->pos4, pd, sizeof(double) * 4 * (size_t)n, s, st));
  SDTFE_RC(upload(t->tets, tetra, sizeof(int) * 8 * (size_t)nt, s, st));

  // the index flag and the incidence counters and cursors start at zero:
  // one array, one clear
  Scratch sc(s);
  const size_t n1 = (size_t)n + 1;
  const size_t tmp_ints = scan_tmp_ints((int64_t)std::max(n + 1, nt));
  SDTFE_RC(sc.reserve(Scratch::size<int>(1 + 2 * n1) + Scratch::size<int>(n1) +
                      Scratch::size<int>(4 * (size_t)nt) + Scratch::size<int>(tmp_ints)));
  int *zeroed = nullptr, *off = nullptr, *inc = nullptr, *tmp = nullptr;
  SDTFE_RC(sc.get(&zeroed, 1 + 2 * n1));
  SDTFE_RC(sc.get(&off, n1));
  SDTFE_RC(sc.get(&inc, 4 * (size_t)nt));
  SDTFE_RC(sc.get(&tmp, tmp_ints));
  int *const bad = zeroed, *const deg = zeroed + 1, *const cursor = zeroed + 1 + n1;
  SDTFE_HIP(hipMemsetAsync(zeroed, 0, sizeof(int) * (1 + 2 * n1), s));
  if (nt > 0) validate_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(t->tets, nt, n, bad);
  int bad_local = 0;
  int *bad_h = readback_words(st, &bad_local);
  SDTFE_HIP(hipMemcpyAsync(bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  SDTFE_HIP(hipStreamSynchronize(s));
  if (*bad_h) return fail(SDTFE_ETOPO, "tetra_data holds an index out of range");

  // volumes and the vertex -> tet incidence (CSR)
  if (nt > 0)
    volume_degree_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(
        reinterpret_cast<const double4 *>(t->pos4), t->tets, nt, t->vol, deg);
  SDTFE_RC(scan_exclusive(deg, off, (int64_t)n + 1, tmp, nullptr, s));
  if (nt > 0)
    fill_incidence_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(t->tets, nt, off, cursor, inc);
  const float m4 = t->p_mass * 4;  // a float product, src/triangulate.c:224
  if (n > 0 && !t->given_field)
    vertex_density_kernel<<<blocks_for(64 * (int64_t)n, 256), 256, 0, s>>>(
        reinterpret_cast<double4 *>(t->pos4), n, off, inc, t->vol, m4);

  // centre and half-extent of the particles' bounding box
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], pd[4 * (size_t)i + k]);
      hi[k] = std::max(hi[k], pd[4 * (size_t)i + k]);
    }
  t->span = 0;
  for (int k = 0; k < 3; ++k) {
    t->org[k] = n > 0 ? 0.5 * (lo[k] + hi[k]) : 0.0;
    if (n > 0) t->span = std::max(t->span, 0.5 * (hi[k] - lo[k]));
  }
  t->z_min = n > 0 ? lo[2] : 0.0;
  t->z_max = n > 0 ? hi[2] : 0.0;

  if (nt > 0) {
    gradient_record_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(
        reinterpret_cast<const double4 *>(t->pos4), t->tets, nt, t->nabla, t->rec64, t->rec32);
    face_record_kernel<<<blocks_for(4 * (int64_t)nt, 256), 256, 0, s>>>(
        reinterpret_cast<const double4 *>(t->pos4), t->tets, nt, t->facerec);
  }
  if (frec64)
    face_record64_kernel<<<blocks_for(4 * (int64_t)nt, 256), 256, 0, s>>>(
        reinterpret_cast<const double4 *>(t->pos4), t->tets, t->nabla, nt, t->frec64);
  // the entry grid follows on the same stream and ends with its sync
  SDTFE_HIP(hipGetLastError());
  return 0;
}

int build_entry_grid(sdtfe_tess *t, Staging *st) {
  const int nt = t->nt;
  hipStream_t s = t->stream;
  Scratch sc(s);
  unsigned char *mask = nullptr;
  int *cnt = nullptr, *off = nullptr, *tmp = nullptr;
  const unsigned nblk = blocks_for(nt, 256);
  double *part = nullptr, *box = nullptr;
  SDTFE_RC(sc.reserve(Scratch::size<unsigned char>((size_t)nt) +
                      2 * Scratch::size<int>((size_t)nt) + Scratch::size<int>(scan_tmp_ints(nt)) +
                      Scratch::size<double>(4 * (size_t)nblk) + Scratch::size<double>(5)));
  SDTFE_RC(sc.get(&mask, (size_t)nt));
  SDTFE_RC(sc.get(&cnt, (size_t)nt));
  SDTFE_RC(sc.get(&off, (size_t)nt));
  SDTFE_RC(sc.get(&tmp, scan_tmp_ints(nt)));
  SDTFE_RC(sc.get(&part, 4 * (size_t)nblk));
  SDTFE_RC(sc.get(&box, 5));
  // the forward hull's triangle count and bounding box, read back together
  int n_tri = 0;
  double bb[4] = {0, 0, 0, 0};
  if (nt > 0) {
    hull_mask_kernel<<<nblk, 256, 0, s>>>(reinterpret_cast<const double4 *>(t->pos4), t->tets,
                                          nt, mask, cnt, part);
    SDTFE_RC(scan_exclusive(cnt, off, nt, tmp, nullptr, s));
    hull_box_kernel<<<1, 256, 0, s>>>(part, (int)nblk, scan_total_device(tmp, nt), box);
    SDTFE_HIP(hipGetLastError());
    double local[5];
    double *h = readback_words(st, local);
    SDTFE_HIP(hipMemcpyAsync(h, box, sizeof local, hipMemcpyDeviceToHost, s));
    SDTFE_HIP(hipStreamSynchronize(s));
    n_tri = (int)h[4];
    if (n_tri > 0)
      for (int j = 0; j < 4; ++j) bb[j] = h[j];
  }
  t->n_tri = n_tri;
  auto layout = [&](Carve &c) {
    c.take(&t->tri_xy, 6 * (size_t)n_tri);
    c.take(&t->tri_tet, (size_t)n_tri);
    c.take(&t->tri_rid, (size_t)n_tri);
    c.take(&t->tri_abc, 3 * (size_t)n_tri);
  };
  Carve sizing;
  layout(sizing);
  SDTFE_RC(dev_block(t, &t->tri_block, sizing.off));
  Carve carve{static_cast<char *>(t->tri_block)};
  layout(carve);
  if (nt > 0)
    emit_tri_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(
        reinterpret_cast<const double4 *>(t->pos4), t->tets, nt, mask, off, t->tri_xy,
        t->tri_tet, t->tri_rid, t->tri_abc);

  EntryGrid g{};
  g.tri_xy = t->tri_xy;
  g.tri_tet = t->tri_tet;
  g.tri_rid = t->tri_rid;
  g.n_tri = n_tri;
  g.nx = g.ny = 1;
  if (n_tri > 0) {
    const double w = bb[2] - bb[0], h = bb[3] - bb[1];
    // about two cells per triangle, square cells, at most 4096 per side
    const double area = w * h;
    if (area > 0) {
      const double cell = std::sqrt(area / (2.0 * n_tri));
      g.nx = (int)std::min(4096.0, std::max(1.0, std::ceil(w / cell)));
      g.ny = (int)std::min(4096.0, std::max(1.0, std::ceil(h / cell)));
    }
  }
  g.x0 = bb[0]; g.y0 = bb[1]; g.x1 = bb[2]; g.y1 = bb[3];
  g.sx = (g.x1 > g.x0) ? g.nx / (g.x1 - g.x0) : 0.0;
  g.sy = (g.y1 > g.y0) ? g.ny / (g.y1 - g.y0) : 0.0;
  const int n_cell = g.nx * g.ny;
  const size_t c1 = (size_t)n_cell + 1;
  // per-cell counts and cursors start at zero: one array, one clear
  int *ccnt = nullptr, *ctmp = nullptr;
  SDTFE_RC(sc.reserve(Scratch::size<int>(2 * c1) + Scratch::size<int>(scan_tmp_ints(n_cell + 1))));
  SDTFE_RC(sc.get(&ccnt, 2 * c1));
  SDTFE_RC(sc.get(&ctmp, scan_tmp_ints(n_cell + 1)));
  int *const ccur = ccnt + c1;
  SDTFE_RC(dev_alloc(t, &t->cell_off, c1));
  SDTFE_HIP(hipMemsetAsync(ccnt, 0, sizeof(int) * 2 * c1, s));
  if (n_tri > 0)
    cell_bin_kernel<false><<<blocks_for(n_tri, 256), 256, 0, s>>>(g, ccnt, nullptr, nullptr);
  int entries_local = 0;
  int *entries_h = readback_words(st, &entries_local);
  SDTFE_RC(scan_exclusive(ccnt, t->cell_off, (int64_t)n_cell + 1, ctmp, entries_h, s));
  const int n_entries = *entries_h;
  SDTFE_RC(dev_alloc(t, &t->cell_tri, (size_t)n_entries));
  g.cell_off = t->cell_off;
  if (n_tri > 0)
    cell_bin_kernel<true><<<blocks_for(n_tri, 256), 256, 0, s>>>(g, ccur, t->cell_off,
                                                                 t->cell_tri);
  g.cell_tri = t->cell_tri;
  t->grid = g;
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipStreamSynchronize(s));
  return 0;
}

// ------------------------------------------- forward hull of host arrays
// find_forward_hull and copy_fh_triangles take host arrays and no
// tessellation handle: the tets and particles go up on a stream of their
// own, the same kernels as a tessellation's setup run, and the results come
// back.  Indices are checked first (SDTFE_ETOPO).
namespace {

struct HostHull {
  hipStream_t s = nullptr;
  double *pos = nullptr;
  int *tets = nullptr;
  ~HostHull() {
    if (!s) return;
    dev_free(pos, s);
    dev_free(tets, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  int init(const double *pd, int n, const int *tetra, int nt) {
    SDTFE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (dev_malloc((void **)&pos, sizeof(double) * 4 * (size_t)(n > 0 ? n : 1), s) !=
            hipSuccess ||
        dev_malloc((void **)&tets, sizeof(int) * 8 * (size_t)(nt > 0 ? nt : 1), s) != hipSuccess)
      return fail(SDTFE_ENOMEM, "forward hull: device allocation failed");
    if (n > 0)
      SDTFE_HIP(hipMemcpyAsync(pos, pd, sizeof(double) * 4 * (size_t)n, hipMemcpyHostToDevice, s));
    if (nt > 0)
      SDTFE_HIP(
          hipMemcpyAsync(tets, tetra, sizeof(int) * 8 * (size_t)nt, hipMemcpyHostToDevice, s));
    Scratch sc(s);
    int *bad = nullptr;
    SDTFE_RC(sc.get(&bad, 1));
    SDTFE_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    if (nt > 0) validate_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(tets, nt, n, bad);
    int bad_h = 0;
    SDTFE_HIP(hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    SDTFE_HIP(hipStreamSynchronize(s));
    if (bad_h) return fail(SDTFE_ETOPO, "tetra_data holds an index out of range");
    return 0;
  }
};

}  // namespace

int hull_mask_host(const double *pd, int n, const int *tetra, int nt, unsigned char *mask) {
  if (nt <= 0) return 0;
  HostHull h;
  SDTFE_RC(h.init(pd, n, tetra, nt));
  Scratch sc(h.s);
  unsigned char *dmask = nullptr;
  int *cnt = nullptr;
  SDTFE_RC(sc.get(&dmask, (size_t)nt));
  SDTFE_RC(sc.get(&cnt, (size_t)nt));
  hull_mask_kernel<<<blocks_for(nt, 256), 256, 0, h.s>>>(reinterpret_cast<const double4 *>(h.pos),
                                                         h.tets, nt, dmask, cnt, nullptr);
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipMemcpyAsync(mask, dmask, (size_t)nt, hipMemcpyDeviceToHost, h.s));
  SDTFE_HIP(hipStreamSynchronize(h.s));
  return 0;
}

int hull_rows_host(const double *pd, int n, const int *tetra, int nt, const unsigned char *mask,
                   int *abc, int *tet, int n_faces) {
  if (nt <= 0 || n_faces <= 0) return 0;
  // each tet's first row: the count of marked faces before it
  std::vector<int> off_h((size_t)nt);
  int64_t total = 0;
  for (int t = 0; t < nt; ++t) {
    off_h[t] = (int)total;
    total += __builtin_popcount(mask[t] & 15u);
  }
  if (total != n_faces) return fail(SDTFE_EINVAL, "forward hull: face count mismatch");
  HostHull h;
  SDTFE_RC(h.init(pd, n, tetra, nt));
  Scratch sc(h.s);
  unsigned char *dmask = nullptr;
  int *off = nullptr, *dtet = nullptr, *drid = nullptr, *dabc = nullptr;
  double *dxy = nullptr;
  SDTFE_RC(sc.get(&dmask, (size_t)nt));
  SDTFE_RC(sc.get(&off, (size_t)nt));
  SDTFE_RC(sc.get(&dtet, (size_t)n_faces));
  SDTFE_RC(sc.get(&drid, (size_t)n_faces));
  SDTFE_RC(sc.get(&dabc, 3 * (size_t)n_faces));
  SDTFE_RC(sc.get(&dxy, 6 * (size_t)n_faces));
  SDTFE_HIP(hipMemcpyAsync(dmask, mask, (size_t)nt, hipMemcpyHostToDevice, h.s));
  SDTFE_HIP(hipMemcpyAsync(off, off_h.data(), sizeof(int) * (size_t)nt, hipMemcpyHostToDevice,
                           h.s));
  emit_tri_kernel<<<blocks_for(nt, 256), 256, 0, h.s>>>(reinterpret_cast<const double4 *>(h.pos),
                                                        h.tets, nt, dmask, off, dxy, dtet, drid,
                                                        dabc);
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipMemcpyAsync(abc, dabc, sizeof(int) * 3 * (size_t)n_faces, hipMemcpyDeviceToHost,
                           h.s));
  SDTFE_HIP(hipMemcpyAsync(tet, dtet, sizeof(int) * (size_t)n_faces, hipMemcpyDeviceToHost, h.s));
  SDTFE_HIP(hipStreamSynchronize(h.s));
  return 0;
}

}  // namespace sdtfe
