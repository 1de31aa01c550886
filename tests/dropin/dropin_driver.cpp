// dropin_driver.cpp -- test driver that calls the reference's C++ entry
// points (mover_PC, interpP2G, mover_interp, particle_sort) through the
// drop-in library, on host structs laid out as the reference's, in the
// order the reference's CPU driver uses (src/sputniPIC_CPU.cpp:219-256).
//
// Input file (native endian, written by tests/test_gpu_dropin.py):
//   int32  nxc ny nzc ns niter nops strict resident sort_cps perx pery perz
//   float64 Lx Ly Lz dt c
//   int32  ops[nops]        1 mover_PC, 2 zero+interpP2G, 3 zero+mover_interp,
//                           4 particle_sort  (each over all species),
//                           5 zero+interpP2G with one thread per species at
//                             once, as the reference's OpenMP team calls it
//                             (src/sputniPIC_CPU.cpp:250-256)
//   FP     fields[6][Nn]    Ex Ey Ez Bxn Byn Bzn
//   per species: int64 n; float64 qom; FP arrays x y z u v w q [n]
// Output file: per species FP x y z u v w q [n], then FP moments[10][Nn].
// FP is the precision the driver is built with (float, or double with
// -DSPIC_REF_FP64).  The grid is built the way setGrid builds it
// (src/Grid.cpp:5-62); the drop-in reads the *_flat arrays, so the chained
// pointer views are left unset.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "spic.h"
#include "sputnipic/sputnipic.h"

static void die(const char *m) {
  std::fprintf(stderr, "dropin_driver: %s\n", m);
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

int main(int argc, char **argv) {
  if (argc != 3) die("usage: dropin_driver <in> <out>");
  FILE *in = std::fopen(argv[1], "rb");
  if (!in) die("cannot open input");
  int hdr[12];
  rd(in, hdr, 12);
  const int nxc = hdr[0], nyc = hdr[1], nzc = hdr[2], ns = hdr[3], niter = hdr[4],
            nops = hdr[5], strict = hdr[6], resident = hdr[7], sort_cps = hdr[8];
  double dh[5];
  rd(in, dh, 5);
  std::vector<int> ops(nops);
  rd(in, ops.data(), nops);

  parameters param;
  param.dt = dh[3];
  param.c = dh[4];
  param.NiterMover = niter;
  param.n_sub_cycles = 1;
  param.PERIODICX = hdr[9] != 0;
  param.PERIODICY = hdr[10] != 0;
  param.PERIODICZ = hdr[11] != 0;
  param.sort_cps = sort_cps;
  param.ns = ns;

  grid grd;
  std::memset(&grd, 0, sizeof(grd));
  grd.nxc = nxc + 2; grd.nyc = nyc + 2; grd.nzc = nzc + 2;
  grd.nxn = grd.nxc + 1; grd.nyn = grd.nyc + 1; grd.nzn = grd.nzc + 1;
  grd.dx = dh[0] / nxc; grd.dy = dh[1] / nyc; grd.dz = dh[2] / nzc;
  grd.invVOL = (FPfield)1.0 / (grd.dx * grd.dy * grd.dz);
  grd.invdx = (FPfield)1.0 / grd.dx;
  grd.invdy = (FPfield)1.0 / grd.dy;
  grd.invdz = (FPfield)1.0 / grd.dz;
  grd.xStart = 0.0; grd.xEnd = dh[0];
  grd.yStart = 0.0; grd.yEnd = dh[1];
  grd.zStart = 0.0; grd.zEnd = dh[2];
  grd.Lx = dh[0]; grd.Ly = dh[1]; grd.Lz = dh[2];
  grd.PERIODICX = param.PERIODICX;
  grd.PERIODICY = param.PERIODICY;
  grd.PERIODICZ = param.PERIODICZ;
  const size_t nn = (size_t)grd.nxn * grd.nyn * grd.nzn;
  std::vector<FPfield> XN(nn), YN(nn), ZN(nn);
  for (int i = 0; i < grd.nxn; i++)
    for (int j = 0; j < grd.nyn; j++)
      for (int k = 0; k < grd.nzn; k++) {
        const size_t n = ((size_t)i * grd.nyn + j) * grd.nzn + k;
        XN[n] = (FPfield)(grd.xStart + (i - 1) * grd.dx);
        YN[n] = (FPfield)(grd.yStart + (j - 1) * grd.dy);
        ZN[n] = (FPfield)(grd.zStart + (k - 1) * grd.dz);
      }
  grd.XN_flat = XN.data(); grd.YN_flat = YN.data(); grd.ZN_flat = ZN.data();

  std::vector<FPfield> F(6 * nn);
  rd(in, F.data(), 6 * nn);
  EMfield field;
  std::memset(&field, 0, sizeof(field));
  field.Ex_flat = &F[0 * nn]; field.Ey_flat = &F[1 * nn]; field.Ez_flat = &F[2 * nn];
  field.Bxn_flat = &F[3 * nn]; field.Byn_flat = &F[4 * nn]; field.Bzn_flat = &F[5 * nn];

  std::vector<particles> part(ns);
  std::vector<std::vector<FPpart>> store(ns);
  std::vector<interpDensSpecies> ids(ns);
  std::vector<std::vector<FPinterp>> mstore(ns);
  for (int is = 0; is < ns; is++) {
    long n;
    double qom;
    rd(in, &n, 1);
    rd(in, &qom, 1);
    std::memset(&part[is], 0, sizeof(particles));
    part[is].species_ID = is;
    part[is].nop = part[is].npmax = n;
    part[is].NiterMover = niter;
    part[is].n_sub_cycles = 1;
    part[is].qom = (FPpart)qom;
    store[is].resize(7 * (size_t)n);
    rd(in, store[is].data(), 7 * (size_t)n);
    FPpart *b = store[is].data();
    part[is].x = b; part[is].y = b + n; part[is].z = b + 2 * n;
    part[is].u = b + 3 * n; part[is].v = b + 4 * n; part[is].w = b + 5 * n;
    part[is].q = (FPinterp *)(b + 6 * n);
    mstore[is].assign(10 * nn, (FPinterp)0);
    std::memset(&ids[is], 0, sizeof(interpDensSpecies));
    FPinterp *mb = mstore[is].data();
    FPinterp **dst[10] = {&ids[is].rhon_flat, &ids[is].Jx_flat,  &ids[is].Jy_flat,
                          &ids[is].Jz_flat,   &ids[is].pxx_flat, &ids[is].pxy_flat,
                          &ids[is].pxz_flat,  &ids[is].pyy_flat, &ids[is].pyz_flat,
                          &ids[is].pzz_flat};
    for (int m = 0; m < 10; m++) *dst[m] = mb + m * nn;
  }
  std::fclose(in);

  spic_set_strict_fp(strict);
  spic_dropin_set_resident(resident);
  for (int op : ops) {
    if (op == 1) {
      for (int is = 0; is < ns; is++)
        if (mover_PC(&part[is], &field, &grd, &param) != 0) die("mover_PC failed");
    } else if (op == 2 || op == 3) {
      for (int is = 0; is < ns; is++)
        std::fill(mstore[is].begin(), mstore[is].end(), (FPinterp)0);
      for (int is = 0; is < ns; is++) {
        if (op == 2) interpP2G(&part[is], &ids[is], &grd);
        else if (mover_interp(&part[is], &field, &ids[is], &grd, &param) != 0)
          die("mover_interp failed");
        if (spic_dropin_last_status() != 0) die("deposit failed");
      }
    } else if (op == 5) {
      for (int is = 0; is < ns; is++)
        std::fill(mstore[is].begin(), mstore[is].end(), (FPinterp)0);
      std::vector<int> status(ns, -1);
      std::vector<std::thread> team;
      for (int is = 0; is < ns; is++)
        team.emplace_back([&, is]() {
          interpP2G(&part[is], &ids[is], &grd);
          status[is] = spic_dropin_last_status();  // the calling thread's status
        });
      for (auto &t : team) t.join();
      for (int is = 0; is < ns; is++)
        if (status[is] != 0) die("threaded interpP2G failed");
    } else if (op == 4) {
      for (int is = 0; is < ns; is++) {
        particle_sort(&param, &part[is], &grd);
        if (spic_dropin_last_status() != 0) die("particle_sort failed");
      }
    } else {
      die("unknown op");
    }
  }
  if (resident)
    for (int is = 0; is < ns; is++) spic_dropin_sync_to_host(&part[is]);

  FILE *out = std::fopen(argv[2], "wb");
  if (!out) die("cannot open output");
  for (int is = 0; is < ns; is++) wr(out, store[is].data(), store[is].size());
  for (int is = 0; is < ns; is++) wr(out, mstore[is].data(), mstore[is].size());
  std::fclose(out);
  spic_dropin_release_all();
  return 0;
}
