// Runs the reference's forward time loop (core/specfem/solver/time_marching.tpp:
// 33-116) through the C++ operator shim on one case read from a binary file,
// and writes the seismograms.  Test driver of tests/test_shim_loop.py.
//
// usage: reference_loop <case.bin> <seismograms.out>
//
// case.bin (little endian): int32 nspec, ngnod, nnodes, nabs, nfree, nsrc,
// nrec, ntypes, nsbs, nstep; float64 dt; then the arrays of smx_mesh2d,
// smx_sources2d and smx_receivers2d in their declaration order (coord f64,
// knods i32, kappa/mu/rho f32, abs_ispec/abs_type i32, free_ispec/free_type
// i32; x/z f64, angle/f0/tshift/factor f32, stf/kind i32, moment f32; x/z
// f64, types i32).  seismograms.out: float32 [nsig][ntypes][nrec][2].
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "specfempp_amd_shim.hpp"

namespace {

struct Reader {
  std::ifstream in;
  explicit Reader(const char *path) : in(path, std::ios::binary) {
    if (!in)
      throw std::runtime_error(std::string("cannot open ") + path);
  }
  template <typename T> std::vector<T> take(size_t n) {
    std::vector<T> v(n);
    if (n && !in.read(reinterpret_cast<char *>(v.data()), n * sizeof(T)))
      throw std::runtime_error("truncated case file");
    return v;
  }
  template <typename T> T one() { return take<T>(1)[0]; }
};

} // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s case.bin seismograms.out\n", argv[0]);
    return 2;
  }
  try {
    Reader r(argv[1]);
    const auto h = r.take<int>(10);
    const int nspec = h[0], ngnod = h[1], nnodes = h[2], nabs = h[3], nfree = h[4],
              nsrc = h[5], nrec = h[6], ntypes = h[7], nsbs = h[8], nstep = h[9];
    const double dt = r.one<double>();
    const auto coord = r.take<double>(size_t(nnodes) * 2);
    const auto knods = r.take<int>(size_t(nspec) * ngnod);
    const auto kappa = r.take<float>(size_t(nspec) * 25);
    const auto mu = r.take<float>(size_t(nspec) * 25);
    const auto rho = r.take<float>(size_t(nspec) * 25);
    const auto abs_ispec = r.take<int>(nabs), abs_type = r.take<int>(nabs);
    const auto free_ispec = r.take<int>(nfree), free_type = r.take<int>(nfree);
    const auto sx = r.take<double>(nsrc), sz = r.take<double>(nsrc);
    const auto angle = r.take<float>(nsrc), f0 = r.take<float>(nsrc),
               tshift = r.take<float>(nsrc), factor = r.take<float>(nsrc);
    const auto stf = r.take<int>(nsrc), kind = r.take<int>(nsrc);
    const auto moment = r.take<float>(size_t(nsrc) * 3);
    const auto rx = r.take<double>(nrec), rz = r.take<double>(nrec);
    const auto types = r.take<int>(ntypes);

    smx_mesh2d mesh{nspec, ngnod, nnodes, coord.data(), knods.data(),
                    kappa.data(), mu.data(), rho.data(), nabs,
                    abs_ispec.data(), abs_type.data(), nfree,
                    free_ispec.data(), free_type.data()};
    smx_sources2d src{nsrc, sx.data(), sz.data(), angle.data(), f0.data(),
                      tshift.data(), factor.data(), stf.data(), kind.data(),
                      moment.data()};
    smx_receivers2d rec{nrec, rx.data(), rz.data(), ntypes, types.data(), nsbs};
    smx_ctx *ctx = nullptr;
    specfempp_amd::check(smx_ctx_create2d(&mesh, &src, &rec, dt, nstep, 4, &ctx));

    // the reference's objects, built from the context instead of an assembly
    using namespace specfempp_amd;
    domain_kernels<simulation_field::forward, 5> kernels(ctx);
    newmark scheme(ctx, nstep, nsbs, float(dt));
    time_marching_forward<5> solver(ctx, scheme, kernels);
    solver.run();

    int nsig = 0;
    check(smx_ctx_info(ctx, nullptr, nullptr, &nsig, nullptr));
    std::vector<float> seis(size_t(nsig) * ntypes * nrec * 2);
    check(smx_get_seismograms(ctx, seis.data()));
    check(smx_ctx_destroy(ctx));
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char *>(seis.data()),
              std::streamsize(seis.size() * sizeof(float)));
    std::printf("reference loop: %d steps, %d samples of %d types at %d receivers\n",
                nstep, nsig, ntypes, nrec);
    return out ? 0 : 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "reference_loop: %s\n", e.what());
    return 1;
  }
}
