// adapter.hip -- the reference's tetra_data layout assembled on the GPU from
// Delaunay simplices (what triangulate does with Qhull's facets,
// src/triangulate.c:56-175):
//
//  * orientation: B and C are swapped where (AB x AC) . AD > 0, so every tet
//    has det(AB, AC, AD) <= 0 (:72-99), in the reference's operation order
//    (this file is built with -ffp-contract=off), so the swap decisions are
//    its own;
//  * neighbours: slot 4 + f holds the tet across the face opposite vertex f,
//    -1 on the hull (:111-175).  Faces are matched in a hash table keyed by
//    their sorted vertex triple: the first face of a pair claims a slot, its
//    twin finds it there and links both tets.  A valid tessellation gives
//    each face at most one twin, so the links are the ones any correct
//    method finds; a third face on one triple is reported.
#include "sdtfe_host.h"

namespace sdtfe {
namespace {

constexpr int kEmpty = -1;  // table slot never claimed; a claimed slot holds
                            // the face id (4 * tet + f), a matched one -2 - id

__global__ void orient_kernel(const double *__restrict__ pd, int n, const int4 *__restrict__ simp,
                              int nt, int4 *tetra, int *bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int4 v = simp[t];
  if (v.x < 0 || v.x >= n || v.y < 0 || v.y >= n || v.z < 0 || v.z >= n || v.w < 0 ||
      v.w >= n) {
    atomicOr(bad, 1);
    v = make_int4(0, 0, 0, 0);
  } else {
    const double *a = pd + 4 * (size_t)v.x, *b = pd + 4 * (size_t)v.y;
    const double *c = pd + 4 * (size_t)v.z, *d = pd + 4 * (size_t)v.w;
    double ab[3], ac[3], ad[3];
    for (int k = 0; k < 3; ++k) {
      ab[k] = b[k] - a[k];
      ac[k] = c[k] - a[k];
      ad[k] = d[k] - a[k];
    }
    const double n0 = ab[1] * ac[2] - ab[2] * ac[1];
    const double n1 = ab[2] * ac[0] - ab[0] * ac[2];
    const double n2 = ab[0] * ac[1] - ab[1] * ac[0];
    double temp = 0;
    temp += n0 * ad[0];
    temp += n1 * ad[1];
    temp += n2 * ad[2];
    if (temp > 0) {
      const int s = v.y;
      v.y = v.z;
      v.z = s;
    }
  }
  tetra[2 * (size_t)t] = v;
  tetra[2 * (size_t)t + 1] = make_int4(-1, -1, -1, -1);
}

__global__ void fill_kernel(int *p, int64_t n, int value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// Sorted vertex triple of face f (the vertices other than f) of tet t.
__device__ inline void face_triple(const int4 *tetra, int t, int f, int &a, int &b, int &c) {
  const int4 v = tetra[2 * (size_t)t];
  const int w[4] = {v.x, v.y, v.z, v.w};
  int q[3], nq = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k != f) q[nq++] = w[k];
  // three-element sorting network
  const int x = min(q[0], q[1]), y = max(q[0], q[1]);
  a = min(x, q[2]);
  const int m = max(x, q[2]);
  b = min(y, m);
  c = max(y, m);
}

__global__ void link_kernel(int4 *tetra, int nt, int *table, uint64_t mask, int *bad) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 4 * (int64_t)nt) return;
  const int t = (int)(id >> 2), f = (int)(id & 3);
  int a, b, c;
  face_triple(tetra, t, f, a, b, c);
  uint64_t slot = mix64(((uint64_t)(uint32_t)a << 32 | (uint32_t)b) ^ mix64((uint64_t)c + kGamma));
  for (uint64_t probe = 0; probe <= mask; ++probe, ++slot) {
    int *s = table + (slot & mask);
    int cur = atomicCAS(s, kEmpty, (int)id);
    if (cur == kEmpty) return;  // first of its pair: wait for the twin
    const bool matched = cur <= -2;
    const int other = matched ? -2 - cur : cur;
    int oa, ob, oc;
    face_triple(tetra, other >> 2, other & 3, oa, ob, oc);
    if (oa != a || ob != b || oc != c) continue;  // another face: probe on
    if (matched || atomicCAS(s, cur, -2 - cur) != cur) {
      atomicOr(bad, 2);  // a third face on this triple
      return;
    }
    reinterpret_cast<int *>(tetra + 2 * (size_t)t + 1)[f] = other >> 2;
    reinterpret_cast<int *>(tetra + 2 * (size_t)(other >> 2) + 1)[other & 3] = t;
    return;
  }
  atomicOr(bad, 4);  // table full (not reached: it has room for every face)
}

}  // namespace

int tetra_from_simplices(const double *pd, int n, const int *simp, int nt, int *tetra_out) {
  if (n < 0 || nt < 0) return fail(SDTFE_EINVAL, "negative size");
  if ((n > 0 && !pd) || (nt > 0 && (!simp || !tetra_out)))
    return fail(SDTFE_EINVAL, "null input array");
  if (nt == 0) return 0;
  if (4 * (int64_t)nt > INT_MAX) return fail(SDTFE_EINVAL, "too many tets for 32-bit face ids");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SDTFE_ENODEV, "no HIP device visible");
  hipStream_t s = nullptr;
  SDTFE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  // hash table of at least 6 slots per tet (about 2 faces per tet are distinct)
  uint64_t cap = 1;
  while (cap < 6 * (uint64_t)nt) cap <<= 1;
  double *d_pd = nullptr;
  int *d_simp = nullptr, *d_tetra = nullptr, *d_table = nullptr, *d_bad = nullptr;
  int rc = 0;
  if (dev_malloc((void **)&d_pd, sizeof(double) * 4 * (size_t)std::max(n, 1), s) != hipSuccess ||
      dev_malloc((void **)&d_simp, sizeof(int) * 4 * (size_t)nt, s) != hipSuccess ||
      dev_malloc((void **)&d_tetra, sizeof(int) * 8 * (size_t)nt, s) != hipSuccess ||
      dev_malloc((void **)&d_table, sizeof(int) * cap, s) != hipSuccess ||
      dev_malloc((void **)&d_bad, sizeof(int), s) != hipSuccess)
    rc = fail(SDTFE_ENOMEM, "device allocation for the tet assembly failed");
  int bad = 0;
  if (!rc) rc = upload(d_pd, pd, sizeof(double) * 4 * (size_t)n, s, nullptr);
  if (!rc) rc = upload(d_simp, simp, sizeof(int) * 4 * (size_t)nt, s, nullptr);
  if (!rc) {
    fill_kernel<<<1, 64, 0, s>>>(d_bad, 1, 0);
    orient_kernel<<<(nt + 255) / 256, 256, 0, s>>>(d_pd, n, reinterpret_cast<int4 *>(d_simp), nt,
                                                  reinterpret_cast<int4 *>(d_tetra), d_bad);
    fill_kernel<<<(unsigned)((cap + 255) / 256), 256, 0, s>>>(d_table, (int64_t)cap, kEmpty);
    const int64_t faces = 4 * (int64_t)nt;
    link_kernel<<<(unsigned)((faces + 255) / 256), 256, 0, s>>>(
        reinterpret_cast<int4 *>(d_tetra), nt, d_table, cap - 1, d_bad);
    rc = hip_check(hipGetLastError(), "tet assembly launch");
  }
  if (!rc)
    rc = hip_check(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(tetra_out, d_tetra, sizeof(int) * 8 * (size_t)nt,
                                  hipMemcpyDeviceToHost, s),
                   "copy");
  dev_free(d_pd, s);
  dev_free(d_simp, s);
  dev_free(d_tetra, s);
  dev_free(d_table, s);
  dev_free(d_bad, s);
  const int rs = hip_check(hipStreamSynchronize(s), "sync");
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  if (rs) return rs;
  if (bad & 1) return fail(SDTFE_ETOPO, "simplices hold a vertex index out of range");
  if (bad & 2) return fail(SDTFE_ETOPO, "a face is shared by more than two tets");
  if (bad & 4) return fail(SDTFE_EHIP, "face table overflow");
  return 0;
}

}  // namespace sdtfe
