// Host build of the device walk code (sdtfe_amd/csrc/walk_fp64.hip,
// walk_fp64_core.h): the same functions compiled for the CPU, driven by
// tests/test_walk_host.py.  Test infrastructure; the library never uses it.
// Input: a binary file with records and rays; output: (value, crossed).
// Modes: FP64 walk over Rec64 (default), "face" (the FP64 walk over face
// records built by make_frec64 from the particles and tetra_data rows that
// follow the rays), "mixed" / "mixedwide" (the mixed walk).
#include "walk_fp64.hip"
#include "walk_fp64_face.h"
#include "walk_mixed_core.h"

#include <cstdio>
#include <string>
#include <vector>

// The face table against the general face choice on all 64 sign patterns.
static int check_faces() {
  int bad = 0;
  for (int w = 0; w < sdtfe::kFaceCodes; ++w) {
    int s[6];
    for (int e = 0; e < 6; ++e) s[e] = (w >> e) & 1 ? -1 : 1;
    // sides CB, AC, BA, DC, BD, DA through plucker64's face logic
    sdtfe::Hit h;
    sdtfe::choose_faces(s[0], s[1], s[2], s[3], s[4], s[5], h);
    const int want = h.ok ? (h.enter | h.leave << 2 | 16) : 0;
    if (sdtfe::face_code(w) != want) {
      fprintf(stderr, "w=%d code %d want %d\n", w, sdtfe::face_code(w), want);
      ++bad;
    }
  }
  return bad ? 9 : 0;
}

int main(int argc, char **argv) {
  if (argc == 2 && std::string(argv[1]) == "faces") return check_faces();
  if (argc != 3 && argc != 4) return 2;
  const std::string mode = argc == 4 ? argv[3] : "";
  const bool face = mode == "face";
  const bool mixed = argc == 4 && !face;  // "mixed": 32-bit face-record offsets, "mixedwide": 64-bit
  const bool wide = mixed && mode == "mixedwide";
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 3;
  long long nt = 0, n = 0;
  double hdr[3];
  unsigned long long seed = 0;
  if (fread(&nt, 8, 1, f) != 1 || fread(&n, 8, 1, f) != 1 || fread(hdr, 8, 3, f) != 3 ||
      fread(&seed, 8, 1, f) != 1)
    return 4;
  std::vector<sdtfe::Rec64> rec((size_t)nt);
  std::vector<double> x((size_t)n), y((size_t)n);
  std::vector<unsigned long long> pix((size_t)n), smp((size_t)n);
  std::vector<long long> ent((size_t)n);
  if (fread(rec.data(), sizeof(sdtfe::Rec64), (size_t)nt, f) != (size_t)nt) return 5;
  if (fread(x.data(), 8, (size_t)n, f) != (size_t)n || fread(y.data(), 8, (size_t)n, f) != (size_t)n ||
      fread(pix.data(), 8, (size_t)n, f) != (size_t)n ||
      fread(smp.data(), 8, (size_t)n, f) != (size_t)n ||
      fread(ent.data(), 8, (size_t)n, f) != (size_t)n)
    return 6;
  std::vector<sdtfe::Rec32> rec32;
  std::vector<sdtfe::FaceRec> frec;
  double org_span[6] = {0, 0, 0, 0, 0, 0};  // origin, half-extent, z range
  if (mixed) {
    rec32.resize((size_t)nt);
    frec.resize(4 * (size_t)nt);
    if (fread(rec32.data(), sizeof(sdtfe::Rec32), (size_t)nt, f) != (size_t)nt ||
        fread(frec.data(), sizeof(sdtfe::FaceRec), 4 * (size_t)nt, f) != 4 * (size_t)nt ||
        fread(org_span, 8, 6, f) != 6)
      return 8;
  }
  std::vector<sdtfe::FRec64> frec64;
  if (face) {
    long long np = 0;
    if (fread(&np, 8, 1, f) != 1) return 10;
    std::vector<double4> pos((size_t)np);
    std::vector<int> tets(8 * (size_t)nt);
    if (fread(pos.data(), sizeof(double4), (size_t)np, f) != (size_t)np ||
        fread(tets.data(), 4, 8 * (size_t)nt, f) != 8 * (size_t)nt)
      return 11;
    std::vector<double> nabla(3 * (size_t)nt);
    for (long long t = 0; t < nt; ++t)
      for (int k = 0; k < 3; ++k) nabla[3 * (size_t)t + k] = rec[(size_t)t].g[k];
    frec64.resize(4 * (size_t)nt);
    for (long long i = 0; i < 4 * nt; ++i)
      sdtfe::make_frec64(pos.data(), tets.data(), nabla.data(), (int)(i >> 2), (int)(i & 3),
                         frec64[(size_t)i]);
  }
  fclose(f);
  sdtfe::MapGeom g{};
  g.lz = hdr[0];
  g.dxy = hdr[1];
  g.off_z = hdr[2];
  g.seed = seed;
  g.max_steps = 1 << 24;
  g.org_x = org_span[0];
  g.org_y = org_span[1];
  g.org_z = org_span[2];
  sdtfe_tess tt;
  tt.span = org_span[3];
  tt.org[0] = org_span[0];
  tt.org[1] = org_span[1];
  tt.org[2] = org_span[2];
  tt.n = mixed ? 1 : 0;
  tt.z_min = org_span[4];
  tt.z_max = org_span[5];
  for (long long i = 0; i < n; ++i)
    g.ray_span = std::max(g.ray_span, std::max(std::fabs(x[i] - g.org_x), std::fabs(y[i] - g.org_y)));
  const sdtfe::MixedConsts mc = sdtfe::mixed_consts(&tt, g);
  uint8_t faces[sdtfe::kFaceCodes];
  for (int w = 0; w < sdtfe::kFaceCodes; ++w) faces[w] = (uint8_t)sdtfe::face_code(w);
  long long redo = 0;
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 7;
  for (long long i = 0; i < n; ++i) {
    sdtfe::LaneCounters c;
    int crossed = 0;
    const uint64_t key = sdtfe::sample_key(seed, pix[i], (uint32_t)smp[i]);
    double v = 0.0;
    // the mixed walk hands some rays to the FP64 walk, as the map kernels do
    if (mixed) {
      // the mixed walk, and the FP64 walk from where it hands a ray over,
      // as the map kernels' second pass does
      const auto *r32 = rec32.data();
      const auto *r64 = rec.data();
      const auto *fr = frec.data();
      const int e = (int)ent[i];
      bool resumed = false;
      if (wide)
        v = mc.flt.full_slab ? sdtfe::walk_mixed_resume<true, true>(r32, r64, fr, faces, e, x[i], y[i], g, mc.flt, key, crossed, resumed, c)
                             : sdtfe::walk_mixed_resume<false, true>(r32, r64, fr, faces, e, x[i], y[i], g, mc.flt, key, crossed, resumed, c);
      else
        v = mc.flt.full_slab ? sdtfe::walk_mixed_resume<true, false>(r32, r64, fr, faces, e, x[i], y[i], g, mc.flt, key, crossed, resumed, c)
                             : sdtfe::walk_mixed_resume<false, false>(r32, r64, fr, faces, e, x[i], y[i], g, mc.flt, key, crossed, resumed, c);
      if (resumed) c.fp64_redo++;
    } else if (face) {
      // the face record the ray enters by: the entry tet's face whose three
      // sides are positive (as the forward-hull triangle names it on the
      // GPU); any other face hands the ray to the walk over Rec64 at once
      const int e = (int)ent[i];
      int rid = e < 0 ? -1 : 4 * e;
      for (int fc = 0; e >= 0 && fc < 4; ++fc) {
        const sdtfe::FRec64 &q = frec64[4 * (size_t)e + fc];
        const double p2 = g.lz, p3 = -g.lz * y[i], p4 = g.lz * x[i];
        const double e0 = sdtfe::edge_u(p2, p3, p4, q.p[6], q.p[7], q.p[3], q.p[4]);
        const double e1 = sdtfe::edge_u(p2, p3, p4, q.p[0], q.p[1], q.p[6], q.p[7]);
        const double e2 = sdtfe::edge_u(p2, p3, p4, q.p[3], q.p[4], q.p[0], q.p[1]);
        if (e0 > sdtfe::kZero && e1 > sdtfe::kZero && e2 > sdtfe::kZero) rid = 4 * e + fc;
      }
      v = sdtfe::walk64f(rec.data(), frec64.data(), rid, x[i], y[i], g, key, crossed, c);
    } else {
      v = sdtfe::walk64(rec.data(), (int)ent[i], x[i], y[i], g, key, crossed, c);
    }
    redo += c.fp64_redo;
    const double cc = crossed;
    fwrite(&v, 8, 1, o);
    fwrite(&cc, 8, 1, o);
  }
  fclose(o);
  if (mixed) fprintf(stderr, "fp64_redo %lld full_slab %d\n", redo, mc.flt.full_slab);
  if (face) fprintf(stderr, "face walk\n");
  return 0;
}
