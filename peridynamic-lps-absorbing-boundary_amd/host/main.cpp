// PD2_LPS_ABC - drop-in replacement of the reference solver's executable (src/main.cpp), running the
// time loop on an MI355X through the lps_amd C ABI.
//
//   PD2_LPS_ABC <prefix> <jobname> [--theta-method=gtheta|idw] [--vtk=on|off] [--exact]
//                                  [--no-graph] [--device=N]
//
// Same positional arguments, input files, .energy/.verify outputs (byte format of
// src/main.cpp:694-707, 1106-1123) and exit codes (0 on success or on the usage message, 1 on a FATAL
// condition) as the reference. VTK snapshots (src/main.cpp:689-691, 1091-1103) are written unless
// --vtk=off. The options select the reference's compile-time variants at run time.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "job_io.h"
#include "lps_amd.h"

namespace {

struct Timer {  // utils::timer (src/helpers/utils.h:263-302)
  std::chrono::steady_clock::time_point t0;
  void tic() { t0 = std::chrono::steady_clock::now(); }
  double toc() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

int fatal_read(const std::string &job, const char *ext) {
  std::cerr << "FATAL: Cannot read " << job << ext << std::endl;
  return 1;
}

}  // namespace

int main(int argc, char *argv[]) {
  std::cout << " _____  ____   ___        _     ____  ____" << std::endl
            << "|  _  ||    \\ |_  |      | |   |  _ \\/ ___| " << std::endl
            << "|   __||  |  ||  _|  ___ | |__ | |_) \\___ \\ " << std::endl
            << "|__|   |____/ |___|_|___||____||____/|____/ " << std::endl
            << std::endl
            << "LPS State-Based PD with ABCs (Example 1) - MI355X" << std::endl
            << std::endl;

  std::vector<std::string> pos;
  lps_options opt;
  lps_default_options(&opt);
  bool vtk = true;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) { pos.push_back(a); continue; }
    if (a == "--theta-method=gtheta") opt.theta_method = LPS_THETA_GTHETA;
    else if (a == "--theta-method=idw") opt.theta_method = LPS_THETA_IDW;
    else if (a == "--vtk=on") vtk = true;
    else if (a == "--vtk=off") vtk = false;
    else if (a == "--exact") opt.exact = 1;
    else if (a == "--no-graph") opt.use_graph = 0;
    else if (a.rfind("--device=", 0) == 0) opt.device = std::atoi(a.c_str() + 9);
    else {
      std::cerr << "unknown option " << a << std::endl;
      return 1;
    }
  }
  if (pos.size() != 2) {
    std::cout << "Usage: " << argv[0] << " <prefix> <jobname>" << std::endl;
    return 0;
  }
  const std::string Prefix = pos[0], JobName = pos[1];

  // ---- read input data (src/main.cpp:176-295) ----
  std::cout << "Reading data... " << std::flush;
  std::vector<int32_t> data1;
  std::vector<double> data2;
  if (!lpsio::read_ints(JobName + ".data1", 6, data1)) return fatal_read(JobName, ".data1");
  if (!lpsio::read_doubles(JobName + ".data2", 15, data2)) return fatal_read(JobName, ".data2");
  const int32_t NodeNo = data1[0], StepNo = data1[1], PrenotchNo = data1[2];
  const int32_t constrainsNo = data1[4], maxCloudSize = data1[5];
  if (PrenotchNo > 0) {
    std::cerr << "FATAL: prenotch bond cutting is not supported by this build" << std::endl;
    return 1;
  }
  std::vector<double> nodes, bcvalues, icd, ic, Gu, Gv, Gth;
  std::vector<int32_t> bctypes, iclouds, infnodes;
  if (!lpsio::read_doubles(JobName + ".nodes", 2 * (int64_t)NodeNo, nodes)) return fatal_read(JobName, ".nodes");
  if (!lpsio::read_ints(JobName + ".bctypes", 2 * (int64_t)NodeNo, bctypes)) return fatal_read(JobName, ".bctypes");
  if (!lpsio::read_doubles(JobName + ".bcvalues", 2 * (int64_t)NodeNo, bcvalues)) return fatal_read(JobName, ".bcvalues");
  if (!lpsio::read_doubles(JobName + ".icdvalues", 2 * (int64_t)NodeNo, icd)) return fatal_read(JobName, ".icdvalues");
  if (!lpsio::read_doubles(JobName + ".icvalues", 2 * (int64_t)NodeNo, ic)) return fatal_read(JobName, ".icvalues");
  if (opt.abc_enabled && constrainsNo > 0) {
    const int64_t C = constrainsNo, M = maxCloudSize;
    if (!lpsio::read_doubles(JobName + ".ciNodesU", 2 * C * 4 * M, Gu)) return fatal_read(JobName, ".ciNodesU");
    if (!lpsio::read_doubles(JobName + ".ciNodesV", 2 * C * 4 * M, Gv)) return fatal_read(JobName, ".ciNodesV");
    if (!lpsio::read_rounded_ints(JobName + ".iClouds", C * M, iclouds)) return fatal_read(JobName, ".iClouds");
    if (!lpsio::read_rounded_ints(JobName + ".infNodesIndex", C, infnodes)) return fatal_read(JobName, ".infNodesIndex");
    if (opt.theta_method == LPS_THETA_GTHETA) {
      if (!lpsio::read_doubles(JobName + ".ciNodesTheta", C * 2 * M, Gth)) return fatal_read(JobName, ".ciNodesTheta");
      std::cout << "(G_theta: " << constrainsNo << " x " << 2 * maxCloudSize << ") ";
    }
  }
  std::cout << "finished." << std::endl;

  lps_problem p{};
  p.node_no = NodeNo;
  p.constrains_no = constrainsNo;
  p.max_cloud_size = maxCloudSize;
  p.E = data2[0]; p.nu = data2[1]; p.rho = data2[2]; p.dt = data2[6];
  p.delta = data2[7]; p.dx = data2[8]; p.dy = data2[9];
  p.nodes = nodes.data();
  p.bctypes = bctypes.data();
  p.icvalues = ic.data();
  p.icdvalues = icd.data();
  p.inf_nodes = infnodes.empty() ? nullptr : infnodes.data();
  p.iclouds = iclouds.empty() ? nullptr : iclouds.data();
  p.ci_nodes_u = Gu.empty() ? nullptr : Gu.data();
  p.ci_nodes_v = Gv.empty() ? nullptr : Gv.data();
  p.ci_nodes_theta = Gth.empty() ? nullptr : Gth.data();

  std::cout << std::endl << "=== Material Parameters ===" << std::endl << std::setprecision(10);
  std::cout << "  E     = " << p.E << std::endl << "  Nu    = " << p.nu << std::endl
            << "  Rho   = " << p.rho << std::endl;
  std::cout << std::endl << "=== Discretization ===" << std::endl;
  std::cout << "  NodeNo = " << NodeNo << std::endl << "  StepNo = " << StepNo << std::endl
            << "  delta (R) = " << p.delta << std::endl << "  dx = " << p.dx << ", dy = " << p.dy
            << std::endl << "  dt = " << p.dt << std::endl << "  ABC nodes = " << constrainsNo
            << std::endl << "  maxCloudSize = " << maxCloudSize << std::endl;

  Timer timer;
  timer.tic();
  lps_ctx *ctx = nullptr;
  int32_t rc = lps_create(&p, &opt, &ctx);
  if (rc != LPS_OK) {
    std::cerr << "FATAL: " << lps_global_error() << std::endl;
    return 1;
  }
  std::cout << std::endl << "Setup on the GPU (horizons, m_i, theta history) took " << timer.toc()
            << " s." << std::endl;
  std::cout << "  m_ref      = " << lps_m_ref(ctx) << std::endl;

  // exclusion radius for the VTK output (src/main.cpp:518-530)
  double domainRadius = 0.0;
  for (int32_t i = 0; i < NodeNo; i++) {
    const double r = std::sqrt(nodes[2 * (size_t)i] * nodes[2 * (size_t)i] +
                               nodes[2 * (size_t)i + 1] * nodes[2 * (size_t)i + 1]);
    if (r > domainRadius) domainRadius = r;
  }
  const double output_exclude_r = (opt.abc_enabled && constrainsNo > 0)
                                      ? domainRadius - p.delta - p.dx / 100.0 : -1.0;

  std::vector<double> u(2 * (size_t)NodeNo), v(2 * (size_t)NodeNo), th((size_t)NodeNo);
  if (vtk) {
    lps_fetch(ctx, u.data(), nullptr, nullptr, nullptr, nullptr);
    lpsio::write_vtk_vector(JobName + Prefix + "-0.vtk", nodes, u, "displacement", NodeNo,
                            output_exclude_r);
  }
  std::ofstream energyFile(JobName + Prefix + ".energy", std::ios::out | std::ios::trunc);
  energyFile << "# t  E_total  E_kinetic  E_potential  P_x  P_y  max_u  max_theta" << std::endl;
  std::ofstream verifyFile(JobName + Prefix + ".verify", std::ios::out | std::ios::trunc);
  verifyFile << "# t  E_total  dE/E0  P_x  P_y  max_theta  max_u  max_v  max_a" << std::endl;

  std::cout << std::endl << "Time marching started (" << StepNo << " steps):" << std::endl;
  timer.tic();
  int32_t done = 0;
  while (done < StepNo) {
    const int32_t chunk = std::min<int32_t>(30 - (int32_t)(lps_step_index(ctx) % 30), StepNo - done);
    lps_diag d;
    rc = lps_run(ctx, chunk, &d, 1);
    done += chunk;
    if (rc < 0 && rc != LPS_ERR_NONFINITE) {
      std::cerr << "FATAL: " << lps_last_error(ctx) << std::endl;
      lps_destroy(ctx);
      return 1;
    }
    if (rc == 0) continue;
    const int64_t s = d.step;
    std::cout << s / 30 << "\tt=" << std::setprecision(6) << d.t << "\tE=" << std::setprecision(8)
              << d.energy << "\tK=" << d.energy_k << "\tP=" << d.energy_p
              << "\tdE/E0=" << std::setprecision(4) << d.de_rel
              << "\t|p|=" << std::sqrt(d.px * d.px + d.py * d.py) << "\tuB=" << d.ub_rms
              << "\tvB=" << d.vb_rms << std::setprecision(10) << std::endl;
    if (vtk) {
      lps_fetch(ctx, u.data(), v.data(), nullptr, th.data(), nullptr);
      const std::string ss = std::to_string(s);
      lpsio::write_vtk_vector(JobName + Prefix + "-" + ss + ".vtk", nodes, u, "displacement",
                              NodeNo, output_exclude_r);
      lpsio::write_vtk_vector(JobName + Prefix + "-vel-" + ss + ".vtk", nodes, v, "velocity",
                              NodeNo, output_exclude_r);
      lpsio::write_vtk_scalar(JobName + Prefix + "-theta-" + ss + ".vtk", nodes, th, "dilatation",
                              NodeNo, output_exclude_r);
    }
    energyFile << std::setprecision(16) << d.t << " " << d.energy << " " << d.energy_k << " "
               << d.energy_p << " " << d.px << " " << d.py << " " << d.max_u << " " << d.theta_max
               << " " << d.ub_rms << " " << d.vb_rms << std::endl;
    verifyFile << std::setprecision(16) << d.t << " " << d.energy << " " << d.de_rel << " " << d.px
               << " " << d.py << " " << d.theta_max << " " << d.max_u << " " << d.max_v << " "
               << d.max_a << std::endl;
    if (std::isnan(d.energy) || std::isinf(d.energy)) {
      std::cerr << "FATAL: Energy is NaN/Inf at step " << s << "!" << std::endl;
      lps_destroy(ctx);
      return 1;
    }
    if (std::isnan(d.max_u) || std::isinf(d.max_u)) {
      std::cerr << "FATAL: Displacement is NaN/Inf at step " << s << "!" << std::endl;
      lps_destroy(ctx);
      return 1;
    }
  }
  std::cout << std::endl << "Time marching took " << timer.toc() << " seconds." << std::endl;
  lps_destroy(ctx);
  return 0;
}
