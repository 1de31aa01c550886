// field_driver.cpp -- test driver of the drop-in's field-solve entry points
// (calculateHatDensities, MaxwellImage) on host structs laid out as the
// reference's, with the chained 3-D views newArr3 would build.
//
// Input file (native endian, written by tests/test_gpu_dropin.py):
//   int32  nxc nyc nzc ns perx pery perz
//   float64 Lx Ly Lz dt c th fourpi qom[ns]
//   FP     B[3][Nn]                 Bxn Byn Bzn
//   FP     per species moments[10][Nn] (rhon Jx Jy Jz pxx pxy pxz pyy pyz pzz)
//   FP     rhoc[Nc]                 net rhoc
//   FP     vec[3 (nxn-2)(nyn-2)(nzn-2)]
// Output file: FP rhoh[Nc], Jxh Jyh Jzh [Nn], im[len(vec)].
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "spic.h"
#include "sputnipic/sputnipic.h"

static void die(const char *m) {
  std::fprintf(stderr, "field_driver: %s\n", m);
  std::exit(2);
}
template <class T>
static void rd(FILE *f, T *p, size_t n) {
  if (std::fread(p, sizeof(T), n, f) != n) die("short read");
}
template <class T>
static void wr(FILE *f, const T *p, size_t n) {
  if (std::fwrite(p, sizeof(T), n, f) != n) die("short write");
}

// Chained view over a contiguous block, as newArr3 lays it out.
struct View3 {
  std::vector<FPinterp *> rows;
  std::vector<FPinterp **> planes;
  FPinterp ***make(FPinterp *base, int n0, int n1, int n2) {
    rows.resize((size_t)n0 * n1);
    planes.resize(n0);
    for (int i = 0; i < n0; i++) {
      for (int j = 0; j < n1; j++) rows[(size_t)i * n1 + j] = base + ((size_t)i * n1 + j) * n2;
      planes[i] = &rows[(size_t)i * n1];
    }
    return planes.data();
  }
};

int main(int argc, char **argv) {
  if (argc != 3) die("usage: field_driver <in> <out>");
  FILE *in = std::fopen(argv[1], "rb");
  if (!in) die("cannot open input");
  int hdr[7];
  rd(in, hdr, 7);
  const int nxc = hdr[0], nyc = hdr[1], nzc = hdr[2], ns = hdr[3];
  std::vector<double> dh(7 + ns);
  rd(in, dh.data(), dh.size());
  parameters param;
  param.Lx = dh[0]; param.Ly = dh[1]; param.Lz = dh[2];
  param.dt = dh[3]; param.c = dh[4]; param.th = dh[5]; param.fourpi = dh[6];
  param.ns = ns;
  for (int s = 0; s < ns; s++) param.qom[s] = dh[7 + s];
  param.PERIODICX = hdr[4] != 0;
  param.PERIODICY = hdr[5] != 0;
  param.PERIODICZ = hdr[6] != 0;

  grid grd;
  std::memset(&grd, 0, sizeof(grd));
  grd.nxc = nxc + 2; grd.nyc = nyc + 2; grd.nzc = nzc + 2;
  grd.nxn = grd.nxc + 1; grd.nyn = grd.nyc + 1; grd.nzn = grd.nzc + 1;
  grd.dx = dh[0] / nxc; grd.dy = dh[1] / nyc; grd.dz = dh[2] / nzc;
  grd.invVOL = (FPfield)1.0 / (grd.dx * grd.dy * grd.dz);
  grd.invdx = (FPfield)1.0 / grd.dx;
  grd.invdy = (FPfield)1.0 / grd.dy;
  grd.invdz = (FPfield)1.0 / grd.dz;
  grd.Lx = dh[0]; grd.Ly = dh[1]; grd.Lz = dh[2];
  grd.PERIODICX = param.PERIODICX;
  grd.PERIODICY = param.PERIODICY;
  grd.PERIODICZ = param.PERIODICZ;
  const size_t nn = (size_t)grd.nxn * grd.nyn * grd.nzn;
  const size_t nc = (size_t)grd.nxc * grd.nyc * grd.nzc;
  std::vector<FPfield> XN(nn), YN(nn), ZN(nn);
  for (int i = 0; i < grd.nxn; i++)
    for (int j = 0; j < grd.nyn; j++)
      for (int k = 0; k < grd.nzn; k++) {
        const size_t n = ((size_t)i * grd.nyn + j) * grd.nzn + k;
        XN[n] = (FPfield)((i - 1) * grd.dx);
        YN[n] = (FPfield)((j - 1) * grd.dy);
        ZN[n] = (FPfield)((k - 1) * grd.dz);
      }
  grd.XN_flat = XN.data(); grd.YN_flat = YN.data(); grd.ZN_flat = ZN.data();

  std::vector<FPfield> B(3 * nn);
  rd(in, B.data(), B.size());
  EMfield field;
  std::memset(&field, 0, sizeof(field));
  field.Ex_flat = field.Bxn_flat = &B[0];
  field.Ey_flat = field.Byn_flat = &B[nn];
  field.Ez_flat = field.Bzn_flat = &B[2 * nn];

  std::vector<FPinterp> M(10 * nn * (size_t)ns);
  rd(in, M.data(), M.size());
  std::vector<interpDensSpecies> ids(ns);
  for (int s = 0; s < ns; s++) {
    std::memset(&ids[s], 0, sizeof(interpDensSpecies));
    FPinterp *b = &M[10 * nn * (size_t)s];
    FPinterp **dst[10] = {&ids[s].rhon_flat, &ids[s].Jx_flat,  &ids[s].Jy_flat,
                          &ids[s].Jz_flat,   &ids[s].pxx_flat, &ids[s].pxy_flat,
                          &ids[s].pxz_flat,  &ids[s].pyy_flat, &ids[s].pyz_flat,
                          &ids[s].pzz_flat};
    for (int m = 0; m < 10; m++) *dst[m] = b + m * nn;
  }
  std::vector<FPinterp> rhoc(nc);
  rd(in, rhoc.data(), nc);
  interpDensNet idn;
  std::memset(&idn, 0, sizeof(idn));
  idn.rhoc_flat = rhoc.data();
  const size_t nvec = 3 * (size_t)(grd.nxn - 2) * (grd.nyn - 2) * (grd.nzn - 2);
  std::vector<FPfield> vec(nvec), im(nvec);
  rd(in, vec.data(), nvec);
  std::fclose(in);

  std::vector<FPinterp> rhoh(nc), jx(nn), jy(nn), jz(nn);
  View3 v0, v1, v2, v3;
  interpDens_aux aux;
  aux.rhoh = v0.make(rhoh.data(), grd.nxc, grd.nyc, grd.nzc);
  aux.Jxh = v1.make(jx.data(), grd.nxn, grd.nyn, grd.nzn);
  aux.Jyh = v2.make(jy.data(), grd.nxn, grd.nyn, grd.nzn);
  aux.Jzh = v3.make(jz.data(), grd.nxn, grd.nyn, grd.nzn);

  calculateHatDensities(&aux, &idn, ids.data(), &field, &grd, &param);
  if (spic_dropin_last_status() != 0) die("calculateHatDensities failed");
  MaxwellImage(im.data(), vec.data(), &field, ids.data(), &grd, &param);
  if (spic_dropin_last_status() != 0) die("MaxwellImage failed");

  FILE *out = std::fopen(argv[2], "wb");
  if (!out) die("cannot open output");
  wr(out, rhoh.data(), nc);
  wr(out, jx.data(), nn);
  wr(out, jy.data(), nn);
  wr(out, jz.data(), nn);
  wr(out, im.data(), nvec);
  std::fclose(out);
  spic_dropin_release_all();
  return 0;
}
