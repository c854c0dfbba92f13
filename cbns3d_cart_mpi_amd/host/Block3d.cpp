// Block3d facade over the C ABI; each method cites the reference method whose
// meaning it keeps.
#include "Block3d.h"

#include <array>
#include <cstdio>
#include <fstream>
#include <memory>
#include <iostream>
#include <stdexcept>

#include "output_vtk_xml.h"

Block3d::Block3d(ProcessInfo* proc_info, SimulationParams* sim_pars, const size_type num_x, const size_type num_y,
                 const size_type num_z)
    : proc_info(proc_info), sim_pars(sim_pars), IM(num_x), JM(num_y), KM(num_z) {
  allocate_mem();
}

Block3d::~Block3d() {
  io_wait();
  io_stop_all();
  free_mem();
}

// ---- asynchronous output ----------------------------------------------------

void Block3d::set_async_io(bool on, int threads) {
  io_wait();
  io_stop_all();
  async_io = on;
  if (on) io_start(threads < 1 ? 1 : threads);
}

void Block3d::io_start(int threads) {
  io_stop = false;
  for (int i = 0; i < threads; i++)
    io_workers.emplace_back([this] {
      for (;;) {
        std::function<int()> job;
        {
          std::unique_lock<std::mutex> lk(io_m);
          io_cv.wait(lk, [this] { return io_stop || !io_jobs.empty(); });
          if (io_jobs.empty()) return;  // stopping, nothing left
          job = std::move(io_jobs.front());
          io_jobs.pop_front();
          io_busy++;
        }
        const int r = job();
        std::lock_guard<std::mutex> lk(io_m);
        io_busy--;
        if (r && !io_err) io_err = r;
        if (io_jobs.empty() && io_busy == 0) io_idle_cv.notify_all();
      }
    });
}

void Block3d::io_stop_all() {
  {
    std::lock_guard<std::mutex> lk(io_m);
    io_stop = true;
  }
  io_cv.notify_all();
  for (auto& w : io_workers) w.join();
  io_workers.clear();
}

void Block3d::io_drain() {
  std::unique_lock<std::mutex> lk(io_m);
  io_idle_cv.wait(lk, [this] { return io_jobs.empty() && io_busy == 0; });
}

int Block3d::io_take_error() {
  std::lock_guard<std::mutex> lk(io_m);
  const int e = io_err;
  io_err = 0;
  return e;
}

int Block3d::io_wait() {
  io_drain();
  return io_take_error() ? -1 : 0;
}

void Block3d::io_enqueue(std::function<int()> job) {
  {
    std::lock_guard<std::mutex> lk(io_m);
    io_jobs.push_back(std::move(job));
  }
  io_cv.notify_one();
}

// A snapshot of the current state, shared by every output call until the
// state changes; a new one waits until the writers are done with the old.
int Block3d::io_snapshot() {
  if (snap_epoch == epoch) return 0;
  io_drain();
  status = cbns_snapshot_begin(ctx);
  if (status != CBNS_OK) return -1;
  snap_epoch = epoch;
  return 0;
}

// src/Block3d.cpp:41-135: all device arrays are owned by one context.
void Block3d::allocate_mem() {
  cbns_params p = to_cbns_params(*sim_pars);
  p.IM = IM;
  p.JM = JM;
  p.KM = KM;
  status = cbns_create(&p, proc_info->dims, CBNS_FP64, &ctx);
  if (status != CBNS_OK) {
    std::cerr << "cbns_create failed: " << cbns_strerror(status) << "\n";
    throw std::runtime_error(cbns_strerror(status));
  }
}

void Block3d::free_mem() {
  if (ctx) cbns_destroy(ctx);
  ctx = nullptr;
}

int Block3d::num_blocks() const {
  int nb = 0;
  cbns_num_blocks(ctx, &nb);
  return nb;
}

int Block3d::which(const value_type* Q) const { return Q == &tag_Q_p ? CBNS_QP : CBNS_Q; }

// src/Block3d.cpp:208-235
void Block3d::gen_mesh(const size_type i_begin, const size_type j_begin, const size_type k_begin) {
  value_type tmp = 2.0 * constant::PI;
  dx = tmp / sim_pars->IM;
  dy = tmp / sim_pars->JM;
  dz = tmp / sim_pars->KM;
  tmp /= 2.0;
  x.resize(IM + 1);
  y.resize(JM + 1);
  z.resize(KM + 1);
  for (size_type i = 0; i < IM + 1; i++) x[i] = (i_begin + i) * dx - tmp;
  for (size_type j = 0; j < JM + 1; j++) y[j] = (j_begin + j) * dy - tmp;
  for (size_type k = 0; k < KM + 1; k++) z[k] = (k_begin + k) * dz - tmp;
}

// src/Block3d.cpp:237-300
void Block3d::initial_condition() {
  touch();
  status = cbns_init_tgv(ctx);
}
void Block3d::calc_conservative(value_type*) {}

value_type* Block3d::get_Q() { return &tag_Q; }
value_type* Block3d::get_Q_p() { return &tag_Q_p; }

// File of block b for the per-rank checkpoint name fname (see Block3d.h):
// fname itself for a single block, else fname with its trailing rank digits
// ("%03d" of rank, then ".dat", as src/cbns3d.cpp:132-134 forms it) replaced
// by b's.  False when several blocks are held and fname has no such suffix.
static bool block_file(const std::string& fname, int rank, int nb, int b, std::string* out) {
  if (nb == 1) {
    *out = fname;
    return true;
  }
  char tail[32];
  std::snprintf(tail, sizeof tail, "%03d.dat", rank);
  const std::string suf(tail);
  if (fname.size() < suf.size() || fname.compare(fname.size() - suf.size(), suf.size(), suf) != 0) return false;
  std::snprintf(tail, sizeof tail, "%03d.dat", b);
  *out = fname.substr(0, fname.size() - suf.size()) + tail;
  return true;
}

// The mu sidecar of a checkpoint file (see Block3d.h).
static std::string mu_file(const std::string& fn) { return fn + ".mu"; }

// src/Block3d.cpp:318-357, one file per block.  mu is not in the file and is
// zero afterwards, as in the reference's restart path, unless a block's mu
// sidecar exists: then that block's mu is set from it.
int Block3d::read_bin(const std::string fname) {
  if (async_io && io_wait()) return -1;  // a pending write may be this file
  touch();
  std::vector<double> all((size_t)NEQ * IM * JM * KM);
  double t = 0.0;
  const int nb = num_blocks();
  std::vector<std::string> files(nb);
  for (int b = 0; b < nb; b++) {
    std::string fn;
    if (!block_file(fname, proc_info->rank, nb, b, &fn)) {
      std::cout << "checkpoint name " << fname << " lacks the rank suffix for " << nb << " blocks\n";
      return -1;
    }
    files[b] = fn;
    std::ifstream fh(fn, std::ios::in | std::ios::binary);
    if (!fh) return -1;
    int n[3], o[3];
    cbns_block_info(ctx, b, n, o);
    double tb;
    fh.read((char*)&tb, sizeof tb);
    size_type d[3];
    fh.read((char*)d, sizeof d);
    if (!fh) return -1;
    for (int a = 0; a < 3; a++)
      if (d[a] != (size_type)(n[a] + 2 * NG)) {
        std::cout << "The dimensions do not match.\n" << std::endl;
        return -1;
      }
    std::vector<double> Qg((size_t)NEQ * d[0] * d[1] * d[2]);
    fh.read((char*)Qg.data(), Qg.size() * sizeof(double));
    if (!fh) return -1;
    for (int k = 0; k < n[2]; k++)
      for (int j = 0; j < n[1]; j++)
        for (int i = 0; i < n[0]; i++)
          for (size_type e = 0; e < NEQ; e++)
            all[NEQ * ((o[0] + i) + (size_t)IM * ((o[1] + j) + (size_t)JM * (o[2] + k))) + e] =
                Qg[NEQ * ((i + NG) + (size_t)d[0] * ((j + NG) + (size_t)d[1] * (k + NG))) + e];
    t = tb;
  }
  sim_pars->t_cur = t;
  status = cbns_set_state(ctx, all.data(), t, 0.0);
  if (status != CBNS_OK) return -1;
  for (int b = 0; b < nb; b++) {
    std::ifstream fh(mu_file(files[b]), std::ios::in | std::ios::binary);
    if (!fh) continue;  // no sidecar: mu stays zero
    int n[3], o[3];
    cbns_block_info(ctx, b, n, o);
    double tb;
    size_type d[3];
    fh.read((char*)&tb, sizeof tb);
    fh.read((char*)d, sizeof d);
    if (!fh || tb != t || d[0] != (size_type)n[0] || d[1] != (size_type)n[1] || d[2] != (size_type)n[2]) {
      std::cout << "The mu sidecar " << mu_file(files[b]) << " does not match its checkpoint.\n";
      return -1;
    }
    std::vector<double> mu((size_t)n[0] * n[1] * n[2]);
    fh.read((char*)mu.data(), mu.size() * sizeof(double));
    if (!fh) return -1;
    status = cbns_set_block_mu(ctx, b, mu.data());
    if (status != CBNS_OK) return -1;
  }
  return 0;
}

// One block's checkpoint file (src/Block3d.cpp:964-987: t_cur, the ghosted
// sizes as three size_type, then the ghosted AoS state) and, given mu, its
// sidecar.
static int write_bin_file(const std::string& fn, double t, const int n[3], const std::vector<double>& Qg,
                          const std::vector<double>* mu) {
  const size_type NG = constant::NG;
  const size_type d[3] = {(size_type)(n[0] + 2 * NG), (size_type)(n[1] + 2 * NG), (size_type)(n[2] + 2 * NG)};
  std::ofstream fh(fn, std::ios::out | std::ios::binary);
  if (!fh) return -1;
  fh.write((const char*)&t, sizeof(value_type));
  fh.write((const char*)d, sizeof d);
  fh.write((const char*)Qg.data(), Qg.size() * sizeof(double));
  if (!fh) return -1;
  if (mu) {
    std::ofstream fm(mu_file(fn), std::ios::out | std::ios::binary);
    const size_type dm[3] = {(size_type)n[0], (size_type)n[1], (size_type)n[2]};
    fm.write((const char*)&t, sizeof(value_type));
    fm.write((const char*)dm, sizeof dm);
    fm.write((const char*)mu->data(), mu->size() * sizeof(double));
    if (!fm) return -1;
  }
  return 0;
}

// src/Block3d.cpp:964-987, one file per block (ghosts included).
int Block3d::output_bin(const std::string fname) {
  const int nb = num_blocks();
  if (async_io && io_take_error()) return -1;
  std::vector<std::string> files(nb);
  for (int b = 0; b < nb; b++)
    if (!block_file(fname, proc_info->rank, nb, b, &files[b])) {
      std::cout << "checkpoint name " << fname << " lacks the rank suffix for " << nb << " blocks\n";
      return -1;
    }
  const double t = sim_pars->t_cur;
  if (async_io && io_snapshot()) return -1;
  for (int b = 0; b < nb; b++) {
    int n[3], o[3];
    cbns_block_info(ctx, b, n, o);
    const size_t cells = (size_t)(n[0] + 2 * NG) * (n[1] + 2 * NG) * (n[2] + 2 * NG);
    // mu is not in the snapshot: read now (a stream synchronisation), written later
    std::shared_ptr<std::vector<double>> mu;
    if (mu_sidecar) {
      mu = std::make_shared<std::vector<double>>((size_t)n[0] * n[1] * n[2]);
      if (cbns_get_block_mu(ctx, b, mu->data()) != CBNS_OK) return -1;
    }
    if (async_io) {
      const std::string fn = files[b];
      const std::array<int, 3> nn = {n[0], n[1], n[2]};
      cbns_ctx* cx = ctx;
      io_enqueue([cx, fn, b, t, nn, cells, mu]() {
        std::vector<double> Qg(NEQ * cells);
        if (cbns_snapshot_block_state(cx, b, Qg.data()) != CBNS_OK) return -1;
        return write_bin_file(fn, t, nn.data(), Qg, mu.get());
      });
      continue;
    }
    std::vector<double> Qg(NEQ * cells);
    if (cbns_get_block_state(ctx, b, CBNS_Q, Qg.data()) != CBNS_OK) return -1;
    if (write_bin_file(files[b], t, n, Qg, mu.get())) return -1;
  }
  return 0;
}

void boundary_condition(Block3d* blk, value_type* Q) {
  blk->touch();
  blk->status = cbns_boundary_condition(blk->ctx, blk->which(Q));
}

void Block3d::calc_primitive(const value_type* Q) {
  touch();
  status = cbns_calc_primitive(ctx, which(Q));
}
void Block3d::init_Q_p() {
  touch();
  status = cbns_init_Q_p(ctx);
}

// src/Block3d.cpp:396-469 for this rank's block
int Block3d::output_vtr(const std::string fname) { return output_vtr(fname, proc_info->rank); }

// The .vtr file of one block (src/Block3d.cpp:396-469): rho u v w p of the
// interior as "%e " ASCII, five per line, then the edge coordinates of the
// block (x, y, z of the whole domain from its origin o).
static int write_vtr_file(const std::string& fname, const int n[3], const int o[3], const std::vector<double> (&f)[5],
                          const std::vector<value_type>* const (&xyz)[3]) {
  const size_type NG = constant::NG;
  const size_t G0 = n[0] + 2 * NG, G1 = n[1] + 2 * NG;
  FILE* fh = std::fopen(fname.c_str(), "w");
  if (!fh) return -1;
  std::fprintf(fh, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fh, "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fh, "  <RectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", n[0], n[1], n[2]);
  std::fprintf(fh, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", n[0], n[1], n[2]);
  std::fprintf(fh, "    <CellData>\n");
  const char* names[5] = {"rho", "u", "v", "w", "p"};
  for (int a = 0; a < 5; a++) {
    std::fprintf(fh, "      <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", names[a]);
    std::fprintf(fh, "        ");
    size_type cnt = 0;
    for (int k = 0; k < n[2]; k++)
      for (int j = 0; j < n[1]; j++)
        for (int i = 0; i < n[0]; i++) {
          std::fprintf(fh, "%e ", f[a][(i + NG) + G0 * ((j + NG) + G1 * (k + NG))]);
          if (4 == cnt++ % 5) std::fprintf(fh, "\n        ");
        }
    if (0 != cnt % 5) std::fprintf(fh, "\n");
    std::fprintf(fh, "      </DataArray>\n");
  }
  std::fprintf(fh, "    </CellData>\n");
  std::fprintf(fh, "    <Coordinates>\n");
  const char* cn[3] = {"x", "y", "z"};
  for (int a = 0; a < 3; a++) {
    std::fprintf(fh, "      <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", cn[a]);
    std::fprintf(fh, "        ");
    const size_type cnt = n[a] + 1;
    for (size_type i = 0; i < cnt; i++) {
      std::fprintf(fh, "%e ", (*xyz[a])[o[a] + i]);
      if (4 == i % 5) std::fprintf(fh, "\n        ");
    }
    if (0 != cnt % 5) std::fprintf(fh, "\n");
    std::fprintf(fh, "      </DataArray>\n");
  }
  std::fprintf(fh, "    </Coordinates>\n");
  std::fprintf(fh, "    </Piece>\n");
  std::fprintf(fh, "  </RectilinearGrid>\n");
  std::fprintf(fh, "</VTKFile>\n");
  return std::fclose(fh) == 0 ? 0 : -1;
}

// src/Block3d.cpp:396-469 for block b
int Block3d::output_vtr(const std::string fname, int b) {
  int n[3], o[3];
  if (cbns_block_info(ctx, b, n, o) != CBNS_OK) return -1;
  const size_t cells = (size_t)(n[0] + 2 * NG) * (n[1] + 2 * NG) * (n[2] + 2 * NG);
  if (async_io) {
    if (io_take_error() || io_snapshot()) return -1;
    const std::array<int, 3> nn = {n[0], n[1], n[2]}, oo = {o[0], o[1], o[2]};
    cbns_ctx* cx = ctx;
    // the edge coordinates are fixed after gen_mesh
    io_enqueue([this, cx, fname, b, nn, oo, cells]() {
      std::vector<double> f[5];
      for (int a = 0; a < 5; a++) {
        f[a].resize(cells);
        if (cbns_snapshot_block_primitive(cx, b, a, f[a].data()) != CBNS_OK) return -1;
      }
      const std::vector<value_type>* const xyz[3] = {&x, &y, &z};
      return write_vtr_file(fname, nn.data(), oo.data(), f, xyz);
    });
    return 0;
  }
  std::vector<double> f[5];
  for (int a = 0; a < 5; a++) {
    f[a].resize(cells);
    if (cbns_get_block_primitive(ctx, b, a, f[a].data()) != CBNS_OK) return -1;
  }
  const std::vector<value_type>* const xyz[3] = {&x, &y, &z};
  return write_vtr_file(fname, n, o, f, xyz);
}

// src/Block3d.cpp:471-500
int Block3d::output_vtm(const std::string fname, const std::string vtr_fname, const size_type num_blocks) {
  FILE* fh = std::fopen(fname.c_str(), "w");
  if (!fh) return -1;
  std::fprintf(fh, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fh, "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fh, "  <vtkMultiBlockDataSet>\n");
  for (size_type i = 0; i < num_blocks; i++)
    std::fprintf(fh, "    <DataSet index=\"%d\" file=\"%s%.3i.vtr\"/>\n", i, vtr_fname.c_str(), i);
  std::fprintf(fh, "  </vtkMultiBlockDataSet>\n");
  std::fprintf(fh, "</VTKFile>\n");
  std::fclose(fh);
  return 0;
}

// src/Block3d.cpp:502-513
bool Block3d::is_finished() { return sim_pars->t_cur + 1.0e-5 > sim_pars->t_end; }

// src/Block3d.cpp:515-553 followed by the MIN over blocks (src/cbns3d.cpp:189)
value_type Block3d::calc_dt() {
  double dt = 0.0;
  status = cbns_calc_dt(ctx, &dt);
  return dt;
}

void Block3d::rec_riemann_x(const value_type* Q) { status = cbns_rec_riemann(ctx, which(Q), 0); }
void Block3d::rec_riemann_y(const value_type* Q) { status = cbns_rec_riemann(ctx, which(Q), 1); }
void Block3d::rec_riemann_z(const value_type* Q) { status = cbns_rec_riemann(ctx, which(Q), 2); }

void Block3d::calc_viscous_terms() {}
void Block3d::calc_viscous_flux() {}
void Block3d::calc_viscous_flux_contribution() { status = cbns_calc_viscous_flux_contribution(ctx); }

void Block3d::update_rk3(const value_type dt, const size_type stage) {
  touch();
  status = cbns_update_rk3(ctx, dt, stage);
}

int Block3d::step(value_type* dt_out) {
  touch();
  cbns_set_time(ctx, sim_pars->t_cur);
  int fin = 0;
  status = cbns_step(ctx, dt_out, &fin);
  cbns_get_time(ctx, &sim_pars->t_cur);
  return status;
}
