// PD2_LPS_ABC - drop-in replacement of the reference solver's executable (src/main.cpp), running the
// time loop on an MI355X through the lps_amd C ABI.
//
//   PD2_LPS_ABC <prefix> <jobname> [--theta-method=gtheta|idw] [--vtk=on|off|binary] [--exact]
//                                  [--no-graph] [--device=N] [--abc=lsq|learned=<weights.bin>]
//   PD2_LPS_ABC --batch=<manifest> [options]
//   restart options: [--checkpoint=<file>] [--restart=<file>] [--stop-at=<step>]
//
// Same positional arguments, input files, .energy/.verify outputs (byte format of
// src/main.cpp:694-707, 1106-1123) and exit codes (0 on success or on the usage message, 1 on a FATAL
// condition) as the reference. VTK snapshots (src/main.cpp:689-691, 1091-1103) are written unless
// --vtk=off. The options select the reference's compile-time variants at run time.
//
// VTK (no counterpart in the reference for these two): --vtk=binary writes legacy BINARY VTK (the
// same fields unrounded, big-endian) instead of ASCII; --vtk-threads=N (default 3, 0 = inline)
// formats and writes the snapshots on N background threads while the GPU goes on stepping.
//
// --batch runs many jobs in one process (no counterpart in the reference): the manifest holds one
// "<prefix> <jobname>" per line ('#' starts a comment); the jobs, which must share dx, dy, delta and
// the step count, are advanced together (lps_create_batch) and each gets its own outputs.
//
// --abc=learned=<weights.bin> (no counterpart in the reference) adds the learned closure to the
// least-squares ABC: weights.bin holds lps_closure_param_count() float32 parameters then sigma_u,
// sigma_v (little endian; `python -m lps_amd.learned --export-bin` writes it). The closure acts where
// a row's max |u| reaches 1e-4 of the job's largest initial displacement, and a job whose energy
// rises falls back to the least-squares ABC (lps_closure_guard).
//
// Checkpoint and restart (no counterpart in the reference): --stop-at=<step> ends the time loop at
// that step, --checkpoint=<file> writes the state after the last step (lps_checkpoint_save), and
// --restart=<file> resumes from such a file (lps_checkpoint_load) and runs on to StepNo, appending
// its records to the .energy/.verify files of the run it continues. A run split this way writes the
// same files, bit for bit, as the run that was never stopped.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "job_io.h"
#include "lps_amd.h"
#include "vtk_async.h"

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

// One job: its files read with the reference's semantics (src/main.cpp:176-295) and its outputs.
struct Job {
  std::string prefix, name;
  int32_t node_no = 0, step_no = 0, constrains_no = 0, max_cloud = 0;
  std::vector<int32_t> data1;
  std::vector<double> data2, nodes, bcvalues, icd, ic, Gu, Gv, Gth, notches;
  std::vector<int32_t> bctypes, iclouds, infnodes;
  double exclude_r = -1.0;
  std::ofstream energy, verify;
  std::vector<double> u, v, th;

  int read(const lps_options &opt) {
    const std::string &J = name;
    if (!lpsio::read_ints(J + ".data1", 6, data1)) return fatal_read(J, ".data1");
    if (!lpsio::read_doubles(J + ".data2", 15, data2)) return fatal_read(J, ".data2");
    node_no = data1[0];
    step_no = data1[1];
    const int32_t prenotch_no = data1[2];
    constrains_no = data1[4];
    max_cloud = data1[5];
    const int64_t N = node_no;
    // prenotch segments, read only when the count is positive (src/main.cpp:251-254)
    if (prenotch_no > 0 && !lpsio::read_doubles(J + ".prenotch", 4 * (int64_t)prenotch_no, notches))
      return fatal_read(J, ".prenotch");
    if (!lpsio::read_doubles(J + ".nodes", 2 * N, nodes)) return fatal_read(J, ".nodes");
    if (!lpsio::read_ints(J + ".bctypes", 2 * N, bctypes)) return fatal_read(J, ".bctypes");
    if (!lpsio::read_doubles(J + ".bcvalues", 2 * N, bcvalues)) return fatal_read(J, ".bcvalues");
    if (!lpsio::read_doubles(J + ".icdvalues", 2 * N, icd)) return fatal_read(J, ".icdvalues");
    if (!lpsio::read_doubles(J + ".icvalues", 2 * N, ic)) return fatal_read(J, ".icvalues");
    if (opt.abc_enabled && constrains_no > 0) {
      const int64_t C = constrains_no, M = max_cloud;
      if (!lpsio::read_doubles(J + ".ciNodesU", 2 * C * 4 * M, Gu)) return fatal_read(J, ".ciNodesU");
      if (!lpsio::read_doubles(J + ".ciNodesV", 2 * C * 4 * M, Gv)) return fatal_read(J, ".ciNodesV");
      if (!lpsio::read_rounded_ints(J + ".iClouds", C * M, iclouds)) return fatal_read(J, ".iClouds");
      if (!lpsio::read_rounded_ints(J + ".infNodesIndex", C, infnodes))
        return fatal_read(J, ".infNodesIndex");
      if (opt.theta_method == LPS_THETA_GTHETA) {
        if (!lpsio::read_doubles(J + ".ciNodesTheta", C * 2 * M, Gth))
          return fatal_read(J, ".ciNodesTheta");
        std::cout << "(G_theta: " << constrains_no << " x " << 2 * max_cloud << ") ";
      }
    }
    // exclusion radius of the VTK output (src/main.cpp:518-530)
    double domain_r = 0.0;
    for (int32_t i = 0; i < node_no; i++)
      domain_r = std::max(domain_r, std::sqrt(nodes[2 * (size_t)i] * nodes[2 * (size_t)i] +
                                              nodes[2 * (size_t)i + 1] * nodes[2 * (size_t)i + 1]));
    exclude_r = (opt.abc_enabled && constrains_no > 0) ? domain_r - data2[7] - data2[8] / 100.0
                                                       : -1.0;
    u.resize(2 * (size_t)node_no);
    v.resize(2 * (size_t)node_no);
    th.resize((size_t)node_no);
    return 0;
  }

  lps_problem problem() const {
    lps_problem p{};
    p.node_no = node_no;
    p.constrains_no = constrains_no;
    p.max_cloud_size = max_cloud;
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
    p.prenotch_no = (int32_t)(notches.size() / 4);
    p.prenotches = notches.empty() ? nullptr : notches.data();
    return p;
  }

  // append: continue the files of the run a restart resumes (no header)
  void open_outputs(bool append = false) {
    const auto mode = append ? std::ios::out | std::ios::app : std::ios::out | std::ios::trunc;
    energy.open(name + prefix + ".energy", mode);
    if (!append)
      energy << "# t  E_total  E_kinetic  E_potential  P_x  P_y  max_u  max_theta" << std::endl;
    verify.open(name + prefix + ".verify", mode);
    if (!append)
      verify << "# t  E_total  dE/E0  P_x  P_y  max_theta  max_u  max_v  max_a" << std::endl;
  }

  // The three snapshots of step s (src/main.cpp:1090-1103) from copies of u, v, theta, queued on
  // the writer; nodes must outlive the writer's wait.
  void write_vtk(int64_t s, lpsio::AsyncWriter &w, lpsio::VtkFormat fmt) {
    const std::string base = name + prefix + "-", ss = std::to_string(s);
    auto cu = std::make_shared<std::vector<double>>(u);
    auto cv = std::make_shared<std::vector<double>>(v);
    auto ct = std::make_shared<std::vector<double>>(th);
    const int32_t n = node_no;
    const double er = exclude_r;
    const std::vector<double> *nd = &nodes;
    w.submit([=] { return lpsio::write_vtk_vector(base + ss + ".vtk", *nd, *cu, "displacement", n, er, fmt); });
    w.submit([=] { return lpsio::write_vtk_vector(base + "vel-" + ss + ".vtk", *nd, *cv, "velocity", n, er, fmt); });
    w.submit([=] { return lpsio::write_vtk_scalar(base + "theta-" + ss + ".vtk", *nd, *ct, "dilatation", n, er, fmt); });
  }
  // The initial displacement snapshot (src/main.cpp:689-691).
  void write_vtk0(lpsio::AsyncWriter &w, lpsio::VtkFormat fmt) {
    auto cu = std::make_shared<std::vector<double>>(u);
    const std::string path = name + prefix + "-0.vtk";
    const int32_t n = node_no;
    const double er = exclude_r;
    const std::vector<double> *nd = &nodes;
    w.submit([=] { return lpsio::write_vtk_vector(path, *nd, *cu, "displacement", n, er, fmt); });
  }

  // One output record (src/main.cpp:1065-1123); returns 1 on the NaN/Inf FATAL (:1126-1133).
  int write_record(const lps_diag &d, const char *tag) {
    const int64_t s = d.step;
    std::cout << tag << s / 30 << "\tt=" << std::setprecision(6) << d.t << "\tE="
              << std::setprecision(8) << d.energy << "\tK=" << d.energy_k << "\tP=" << d.energy_p
              << "\tdE/E0=" << std::setprecision(4) << d.de_rel
              << "\t|p|=" << std::sqrt(d.px * d.px + d.py * d.py) << "\tuB=" << d.ub_rms
              << "\tvB=" << d.vb_rms << std::setprecision(10) << std::endl;
    energy << std::setprecision(16) << d.t << " " << d.energy << " " << d.energy_k << " "
           << d.energy_p << " " << d.px << " " << d.py << " " << d.max_u << " " << d.theta_max
           << " " << d.ub_rms << " " << d.vb_rms << std::endl;
    verify << std::setprecision(16) << d.t << " " << d.energy << " " << d.de_rel << " " << d.px
           << " " << d.py << " " << d.theta_max << " " << d.max_u << " " << d.max_v << " "
           << d.max_a << std::endl;
    if (std::isnan(d.energy) || std::isinf(d.energy)) {
      std::cerr << "FATAL: " << tag << "Energy is NaN/Inf at step " << s << "!" << std::endl;
      return 1;
    }
    if (std::isnan(d.max_u) || std::isinf(d.max_u)) {
      std::cerr << "FATAL: " << tag << "Displacement is NaN/Inf at step " << s << "!" << std::endl;
      return 1;
    }
    return 0;
  }
};

// The learned closure of --abc=learned=<path>, attached to a context; destroyed with the holder.
struct Learned {
  lps_closure *clo = nullptr;
  ~Learned() { lps_closure_destroy(clo); }
  // 0 on success, else prints the FATAL message and returns 1.
  int attach(lps_ctx *ctx, const std::string &path, int32_t device,
             const std::vector<float> &gate) {
    const int64_t np = lps_closure_param_count();
    std::ifstream f(path, std::ios::binary);
    std::vector<float> w(np + 2);
    if (!f || !f.read(reinterpret_cast<char *>(w.data()), sizeof(float) * w.size()) ||
        f.peek() != std::char_traits<char>::eof()) {
      std::cerr << "FATAL: " << path << " is not a closure weight file (" << np + 2
                << " float32 values)" << std::endl;
      return 1;
    }
    if (lps_closure_create(device, 0, &clo) != LPS_OK ||
        lps_closure_set_params(clo, w.data()) != LPS_OK ||
        lps_closure_set_scales(clo, w[np], w[np + 1]) != LPS_OK) {
      std::cerr << "FATAL: learned closure setup failed: " << lps_closure_error(clo) << std::endl;
      return 1;
    }
    if (lps_closure_attach(ctx, clo) != LPS_OK || lps_closure_set_gate(ctx, gate.data()) != LPS_OK) {
      std::cerr << "FATAL: " << lps_last_error(ctx) << std::endl;
      return 1;
    }
    std::cout << "Learned ABC closure " << path << " attached." << std::endl;
    return 0;
  }
  void report(lps_ctx *ctx, const std::vector<std::string> &names) {
    if (!clo) return;
    std::vector<int64_t> fb(names.size(), -1);
    lps_closure_guard(ctx, -1.0, fb.data());
    for (size_t i = 0; i < names.size(); i++)
      if (fb[i] >= 0)
        std::cout << "[" << names[i] << "] learned closure fell back to the least-squares ABC at "
                  << "step " << fb[i] << " (energy rise)" << std::endl;
  }
};

// --vtk, --vtk-threads
struct Vtk {
  bool on = true;
  lpsio::VtkFormat fmt = lpsio::VtkFormat::Ascii;
  int threads = 3;
};

// Waits for the queued snapshots; a failed write is reported, as the reference does not check.
void finish_vtk(lpsio::AsyncWriter &w) {
  const int failed = w.wait();
  if (failed > 0) std::cerr << "WARNING: " << failed << " VTK file(s) could not be written" << std::endl;
}

// --checkpoint, --restart, --stop-at
struct Restart {
  std::string save, load;
  int64_t stop_at = -1;
  // The step the loop ends at.
  int64_t end(int64_t step_no) const { return stop_at >= 0 ? std::min(stop_at, step_no) : step_no; }
  // 0 on success, else prints the FATAL message and returns 1.
  int resume(lps_ctx *ctx) const {
    if (load.empty()) return 0;
    if (lps_checkpoint_load(ctx, load.c_str()) != LPS_OK) {
      std::cerr << "FATAL: " << lps_last_error(ctx) << std::endl;
      return 1;
    }
    std::cout << "Restarted from " << load << " at step " << lps_step_index(ctx) << "."
              << std::endl;
    return 0;
  }
  int finish(lps_ctx *ctx) const {
    if (save.empty()) return 0;
    if (lps_checkpoint_save(ctx, save.c_str()) != LPS_OK) {
      std::cerr << "FATAL: " << lps_last_error(ctx) << std::endl;
      return 1;
    }
    std::cout << "Checkpoint of step " << lps_step_index(ctx) << " written to " << save << "."
              << std::endl;
    return 0;
  }
};

float gate_of(const std::vector<double> &ic) {
  double m = 0.0;
  for (double x : ic) m = std::max(m, std::fabs(x));
  return (float)(1e-4 * m);
}

int run_single(Job &job, const lps_options &opt, const Vtk &vtk, const std::string &learned,
               const Restart &rs) {
  lpsio::AsyncWriter writer(vtk.on ? vtk.threads : 0);
  std::cout << "Reading data... " << std::flush;
  if (job.read(opt)) return 1;
  std::cout << "finished." << std::endl;
  const lps_problem p = job.problem();
  std::cout << std::endl << "=== Material Parameters ===" << std::endl << std::setprecision(10);
  std::cout << "  E     = " << p.E << std::endl << "  Nu    = " << p.nu << std::endl
            << "  Rho   = " << p.rho << std::endl;
  std::cout << std::endl << "=== Discretization ===" << std::endl;
  std::cout << "  NodeNo = " << job.node_no << std::endl << "  StepNo = " << job.step_no
            << std::endl << "  delta (R) = " << p.delta << std::endl << "  dx = " << p.dx
            << ", dy = " << p.dy << std::endl << "  dt = " << p.dt << std::endl
            << "  ABC nodes = " << job.constrains_no << std::endl
            << "  maxCloudSize = " << job.max_cloud << std::endl;

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
  Learned closure;
  if (!learned.empty() && closure.attach(ctx, learned, opt.device, {gate_of(job.ic)})) {
    lps_destroy(ctx);
    return 1;
  }
  if (rs.resume(ctx)) {
    lps_destroy(ctx);
    return 1;
  }
  if (vtk.on && rs.load.empty()) {
    lps_fetch(ctx, job.u.data(), nullptr, nullptr, nullptr, nullptr);
    job.write_vtk0(writer, vtk.fmt);
  }
  job.open_outputs(!rs.load.empty());
  const int64_t end = rs.end(job.step_no);
  std::cout << std::endl << "Time marching started (" << end - lps_step_index(ctx) << " steps):"
            << std::endl;
  timer.tic();
  while (lps_step_index(ctx) < end) {
    const int64_t s0 = lps_step_index(ctx);
    const int32_t chunk = (int32_t)std::min<int64_t>(30 - s0 % 30, end - s0);
    lps_diag d;
    rc = lps_run(ctx, chunk, &d, 1);
    if (rc < 0 && rc != LPS_ERR_NONFINITE) {
      std::cerr << "FATAL: " << lps_last_error(ctx) << std::endl;
      lps_destroy(ctx);
      return 1;
    }
    if (rc == 0) continue;
    if (vtk.on) {
      lps_fetch(ctx, job.u.data(), job.v.data(), nullptr, job.th.data(), nullptr);
      job.write_vtk(d.step, writer, vtk.fmt);
    }
    if (job.write_record(d, "")) {
      lps_destroy(ctx);
      return 1;
    }
  }
  finish_vtk(writer);
  std::cout << std::endl << "Time marching took " << timer.toc() << " seconds." << std::endl;
  closure.report(ctx, {job.name});
  rc = rs.finish(ctx);
  lps_destroy(ctx);
  return rc;
}

int run_batch(const std::string &manifest, const lps_options &opt, const Vtk &vtk,
              const std::string &learned, const Restart &rs) {
  std::ifstream mf(manifest);
  if (!mf) {
    std::cerr << "FATAL: Cannot read " << manifest << std::endl;
    return 1;
  }
  std::vector<Job> jobs;
  std::string line;
  while (std::getline(mf, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ls(line);
    std::string prefix, name;
    if (!(ls >> prefix >> name)) continue;
    jobs.emplace_back();
    jobs.back().prefix = prefix;
    jobs.back().name = name;
  }
  if (jobs.empty()) {
    std::cerr << "FATAL: " << manifest << " lists no job" << std::endl;
    return 1;
  }
  std::cout << "Reading " << jobs.size() << " jobs... " << std::flush;
  std::vector<lps_problem> probs;
  for (Job &j : jobs) {
    if (j.read(opt)) return 1;
    probs.push_back(j.problem());
    if (j.step_no != jobs[0].step_no) {
      std::cerr << "FATAL: the jobs of a batch need the same StepNo (" << j.name << ")"
                << std::endl;
      return 1;
    }
  }
  std::cout << "finished." << std::endl;
  Timer timer;
  timer.tic();
  lps_ctx *ctx = nullptr;
  int32_t rc = lps_create_batch(probs.data(), (int32_t)probs.size(), &opt, &ctx);
  if (rc != LPS_OK) {
    std::cerr << "FATAL: " << lps_global_error() << std::endl;
    return 1;
  }
  std::cout << "Setup on the GPU took " << timer.toc() << " s: " << lps_describe(ctx) << std::endl;
  const int32_t B = (int32_t)jobs.size();
  Learned closure;
  std::vector<std::string> names;
  for (Job &j : jobs) names.push_back(j.name);
  if (!learned.empty()) {
    std::vector<float> gate;
    for (Job &j : jobs) gate.push_back(gate_of(j.ic));
    if (closure.attach(ctx, learned, opt.device, gate)) {
      lps_destroy(ctx);
      return 1;
    }
  }
  if (rs.resume(ctx)) {
    lps_destroy(ctx);
    return 1;
  }
  lpsio::AsyncWriter writer(vtk.on ? vtk.threads : 0);
  for (int32_t i = 0; i < B; i++) {
    Job &j = jobs[i];
    if (vtk.on && rs.load.empty()) {
      lps_fetch_instance(ctx, i, j.u.data(), nullptr, nullptr, nullptr, nullptr);
      j.write_vtk0(writer, vtk.fmt);
    }
    j.open_outputs(!rs.load.empty());
  }
  const int64_t end = rs.end(jobs[0].step_no);
  std::cout << std::endl << "Time marching started (" << end - lps_step_index(ctx) << " steps, "
            << B << " jobs):" << std::endl;
  timer.tic();
  std::vector<lps_diag> d(B);
  while (lps_step_index(ctx) < end) {
    const int64_t s0 = lps_step_index(ctx);
    const int32_t chunk = (int32_t)std::min<int64_t>(30 - s0 % 30, end - s0);
    rc = lps_run_batch(ctx, chunk, d.data(), 1);
    if (rc < 0 && rc != LPS_ERR_NONFINITE) {
      std::cerr << "FATAL: " << lps_last_error(ctx) << std::endl;
      lps_destroy(ctx);
      return 1;
    }
    if (rc == 0) continue;
    int fatal = 0;
    for (int32_t i = 0; i < B; i++) {
      Job &j = jobs[i];
      if (vtk.on) {
        lps_fetch_instance(ctx, i, j.u.data(), j.v.data(), nullptr, j.th.data(), nullptr);
        j.write_vtk(d[i].step, writer, vtk.fmt);
      }
      const std::string tag = "[" + j.name + "] ";
      fatal |= j.write_record(d[i], tag.c_str());
    }
    if (fatal) {
      lps_destroy(ctx);
      return 1;
    }
  }
  finish_vtk(writer);
  std::cout << std::endl << "Time marching took " << timer.toc() << " seconds." << std::endl;
  closure.report(ctx, names);
  rc = rs.finish(ctx);
  lps_destroy(ctx);
  return rc;
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
  std::string manifest, learned;
  Restart rs;
  lps_options opt;
  lps_default_options(&opt);
  Vtk vtk;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) { pos.push_back(a); continue; }
    if (a == "--theta-method=gtheta") opt.theta_method = LPS_THETA_GTHETA;
    else if (a == "--theta-method=idw") opt.theta_method = LPS_THETA_IDW;
    else if (a == "--vtk=on") vtk.on = true, vtk.fmt = lpsio::VtkFormat::Ascii;
    else if (a == "--vtk=off") vtk.on = false;
    else if (a == "--vtk=binary") vtk.on = true, vtk.fmt = lpsio::VtkFormat::Binary;
    else if (a.rfind("--vtk-threads=", 0) == 0) vtk.threads = std::max(0, std::atoi(a.c_str() + 14));
    else if (a == "--exact") opt.exact = 1;
    else if (a == "--no-graph") opt.use_graph = 0;
    else if (a.rfind("--device=", 0) == 0) opt.device = std::atoi(a.c_str() + 9);
    else if (a.rfind("--batch=", 0) == 0) manifest = a.substr(8);
    else if (a == "--abc=lsq") learned.clear();
    else if (a.rfind("--abc=learned=", 0) == 0) learned = a.substr(14);
    else if (a.rfind("--checkpoint=", 0) == 0) rs.save = a.substr(13);
    else if (a.rfind("--restart=", 0) == 0) rs.load = a.substr(10);
    else if (a.rfind("--stop-at=", 0) == 0) rs.stop_at = std::atoll(a.c_str() + 10);
    else {
      std::cerr << "unknown option " << a << std::endl;
      return 1;
    }
  }
  if (!learned.empty() && opt.exact) {
    std::cerr << "FATAL: the learned closure runs on the fast path only (drop --exact)" << std::endl;
    return 1;
  }
  if (!learned.empty() && !(rs.save.empty() && rs.load.empty())) {
    std::cerr << "FATAL: checkpoints do not hold the learned closure's state (drop --abc=learned)"
              << std::endl;
    return 1;
  }
  if (!manifest.empty()) return run_batch(manifest, opt, vtk, learned, rs);
  if (pos.size() != 2) {
    std::cout << "Usage: " << argv[0] << " <prefix> <jobname>" << std::endl;
    return 0;
  }
  Job job;
  job.prefix = pos[0];
  job.name = pos[1];
  return run_single(job, opt, vtk, learned, rs);
}
