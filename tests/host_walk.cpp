// Host build of the device walk code (sdtfe_amd/csrc/walk_fp64.hip,
// walk_fp64_core.h): the same functions compiled for the CPU, driven by
// tests/test_walk_host.py.  Test infrastructure; the library never uses it.
// Input: a binary file with records and rays; output: (value, crossed).
#include "walk_fp64.hip"
#include "walk_mixed_core.h"

#include <cstdio>
#include <vector>

int main(int argc, char **argv) {
  if (argc != 3 && argc != 4) return 2;
  const bool mixed = argc == 4;
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
  double org_span[4] = {0, 0, 0, 0};
  if (mixed) {
    rec32.resize((size_t)nt);
    if (fread(rec32.data(), sizeof(sdtfe::Rec32), (size_t)nt, f) != (size_t)nt ||
        fread(org_span, 8, 4, f) != 4)
      return 8;
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
  for (long long i = 0; i < n; ++i)
    g.ray_span = std::max(g.ray_span, std::max(std::fabs(x[i] - g.org_x), std::fabs(y[i] - g.org_y)));
  const sdtfe::MixedConsts mc = sdtfe::mixed_consts(&tt, g);
  long long redo = 0;
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 7;
  for (long long i = 0; i < n; ++i) {
    sdtfe::LaneCounters c;
    int crossed = 0;
    const uint64_t key = sdtfe::sample_key(seed, pix[i], (uint32_t)smp[i]);
    const double v =
        mixed ? sdtfe::walk_mixed(rec32.data(), rec.data(), (int)ent[i], x[i], y[i], g, mc.flt,
                                  key, crossed, c)
              : sdtfe::walk64(rec.data(), (int)ent[i], x[i], y[i], g, key, crossed, c);
    redo += c.fp64_redo;
    const double cc = crossed;
    fwrite(&v, 8, 1, o);
    fwrite(&cc, 8, 1, o);
  }
  fclose(o);
  if (mixed) fprintf(stderr, "fp64_redo %lld\n", redo);
  return 0;
}
