// Drives the C++ adapter of include/thcgw.hpp the way CoQui's scf_loop drives its solvers
// (update_w, evaluate, hermitize, rpa_energy) and checks it against the one-call
// thcgw_gw_iteration.  Inputs are raw little-endian files written by
// tests/test_gpu_cpp_adapter.py into <dir>; the adapter's Sigma and W are written back there.
//
//   adapter_check <dir>
//
// Prints one JSON line with the relative differences; exit status 0 when they are <= 1e-12.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "thcgw.hpp"

using thcgw_cxx::cplx;

template <class T>
static std::vector<T> load(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t n = static_cast<size_t>(f.tellg()) / sizeof(T);
  f.seekg(0);
  std::vector<T> v(n);
  f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
  return v;
}

template <class T>
static void save(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

static void ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
static T* dev_from(const std::vector<T>& h) {
  T* p = nullptr;
  ok(hipMalloc(&p, h.size() * sizeof(T)), "hipMalloc");
  ok(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return p;
}

template <class T>
static std::vector<T> host_from(const T* d, size_t n) {
  std::vector<T> h(n);
  ok(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return h;
}

static double reldiff(const std::vector<cplx>& a, const std::vector<cplx>& b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num = std::max(num, std::abs(a[i] - b[i]));
    den = std::max(den, std::abs(b[i]));
  }
  return den > 0 ? num / den : num;
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: adapter_check <dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  try {
    std::vector<double> dimv = load<double>(dir + "/dims.bin");  // kgrid[3], no, nmu, nt, nw_b, nw_f, beta
    thcgw_dims d{};
    for (int i = 0; i < 3; ++i) d.kgrid[i] = static_cast<int>(dimv[i]);
    d.no = static_cast<int>(dimv[3]);
    d.nmu = static_cast<int>(dimv[4]);
    d.ns = 1;
    d.nt = static_cast<int>(dimv[5]);
    d.nw_b = static_cast<int>(dimv[6]);
    d.nw_f = static_cast<int>(dimv[7]);
    d.beta = dimv[8];
    std::vector<double> tau = load<double>(dir + "/tau_mesh_f.bin");
    std::vector<long> wn = load<long>(dir + "/wn_mesh_f.bin");
    std::vector<double> Ttw_ff = load<double>(dir + "/Ttw_ff.bin"), Twt_ff = load<double>(dir + "/Twt_ff.bin");
    std::vector<double> Ttw_bb = load<double>(dir + "/Ttw_bb.bin"), Twt_bb = load<double>(dir + "/Twt_bb.bin");
    std::vector<double> Tb = load<double>(dir + "/T_beta.bin");
    thcgw_ir_tables t{tau.data(), wn.data(), Ttw_ff.data(), Twt_ff.data(), Ttw_bb.data(), Twt_bb.data(), Tb.data()};

    thcgw_cxx::context ctx(d, t);
    const size_t nk = ctx.nk(), M = size_t(d.nmu) * d.nmu;
    const size_t nG = size_t(d.nt) * nk * d.no * d.no, nW = nk * ctx.nt_half() * M;
    std::vector<cplx> G = load<cplx>(dir + "/G.bin"), X = load<cplx>(dir + "/X.bin"), V = load<cplx>(dir + "/V.bin");
    if (G.size() != nG || X.size() != nk * d.nmu * d.no || V.size() != nk * M)
      throw std::runtime_error("input sizes do not match dims");
    cplx* dG = dev_from(G);
    cplx* dX = dev_from(X);
    cplx* dV = dev_from(V);
    cplx *dS = nullptr, *dW = nullptr, *dS2 = nullptr, *dW2 = nullptr;
    ok(hipMalloc(&dS, nG * sizeof(cplx)), "hipMalloc");
    ok(hipMalloc(&dW, nW * sizeof(cplx)), "hipMalloc");
    ok(hipMalloc(&dS2, nG * sizeof(cplx)), "hipMalloc");
    ok(hipMalloc(&dW2, nW * sizeof(cplx)), "hipMalloc");

    // the reference's call sequence: update_w, evaluate, hermitize (scf_driver.cpp:119-128), rpa_energy
    thcgw_cxx::MBState mb{dG, dS, dW};
    thcgw_cxx::THCView thc{dX, dV, d.nmu};
    thcgw_cxx::scr_coulomb_t scr(ctx);
    thcgw_cxx::gw_t gw(ctx);
    scr.update_w(mb, thc, 0);
    gw.evaluate(mb, thc);
    gw.hermitize(mb);
    const double ec = gw.rpa_energy(dG, thc);
    ok(hipDeviceSynchronize(), "sync");

    double ec2 = 0.0;
    thcgw_cxx::check(ctx.get(), thcgw_gw_iteration(ctx.get(), thcgw_cxx::D(dG), thcgw_cxx::D(dX), thcgw_cxx::D(dV),
                                                   thcgw_cxx::D(dS2), thcgw_cxx::D(dW2), &ec2, 1, nullptr));
    std::vector<cplx> S = host_from(dS, nG), W = host_from(dW, nW), S2 = host_from(dS2, nG), W2 = host_from(dW2, nW);
    save(dir + "/Sigma_adapter.bin", S);
    save(dir + "/W_adapter.bin", W);
    const double es = reldiff(S, S2), ew = reldiff(W, W2), ee = std::abs(ec - ec2) / std::abs(ec2);
    std::printf("{\"sigma_rel\": %.3e, \"w_rel\": %.3e, \"ec_rel\": %.3e, \"ec\": %.15e}\n", es, ew, ee, ec);

    // the same sequence with the long-wavelength treatment (gygi_extrplt), when its inputs are given
    std::FILE* probe = std::fopen((dir + "/chi_bar.bin").c_str(), "rb");
    if (probe) {
      std::fclose(probe);
      std::vector<cplx> cb = load<cplx>(dir + "/chi_bar.bin"), ch = load<cplx>(dir + "/chi_head.bin");
      std::vector<double> q = load<double>(dir + "/qpts.bin"), par = load<double>(dir + "/gygi.bin");  // volume, madelung
      cplx* dcb = dev_from(cb);
      cplx* dch = dev_from(ch);
      thcgw_cxx::THCView thc2 = thc;
      thc2.chi_bar_qP = dcb;
      thc2.chi_head_P = dch;
      thc2.qpts_cart = q.data();
      thc2.volume = par[0];
      thc2.madelung = par[1];
      thcgw_cxx::MBState mb2{dG, dS, dW};
      thcgw_cxx::scr_coulomb_t scr2(ctx, nullptr, thcgw_cxx::gygi_extrplt);
      thcgw_cxx::gw_t gw2(ctx, nullptr, thcgw_cxx::gygi_extrplt);
      scr2.update_w(mb2, thc2, 1);
      gw2.evaluate(mb2, thc2);
      gw2.hermitize(mb2);
      ok(hipDeviceSynchronize(), "sync");
      save(dir + "/Sigma_gygi.bin", host_from(dS, nG));
      save(dir + "/eps_head.bin", mb2.eps_inv_head);
      for (cplx* p : {dcb, dch}) (void)hipFree(p);
    }
    // the plain sequence again with the resolvent Dyson form (flag 6 = 2)
    ctx.set_dyson_form(thcgw_cxx::context::dyson_resolvent);
    thcgw_cxx::MBState mb3{dG, dS, dW};
    thcgw_cxx::scr_coulomb_t scr3(ctx);
    thcgw_cxx::gw_t gw3(ctx);
    scr3.update_w(mb3, thc, 0);
    gw3.evaluate(mb3, thc);
    gw3.hermitize(mb3);
    const double ec3 = gw3.rpa_energy(dG, thc);
    ok(hipDeviceSynchronize(), "sync");
    save(dir + "/Sigma_resolvent.bin", host_from(dS, nG));
    save(dir + "/W_resolvent.bin", host_from(dW, nW));
    save(dir + "/ec_resolvent.bin", std::vector<double>{ec3});
    ctx.set_dyson_form(thcgw_cxx::context::dyson_lu);
    for (cplx* p : {dG, dX, dV, dS, dW, dS2, dW2}) (void)hipFree(p);
    return (es <= 1e-12 && ew <= 1e-12 && ee <= 1e-12) ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "adapter_check: %s\n", e.what());
    return 3;
  }
}
