// Block3d facade over the C ABI; each method cites the reference method whose
// meaning it keeps.
#include "Block3d.h"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "output_vtk_xml.h"

Block3d::Block3d(ProcessInfo* proc_info, SimulationParams* sim_pars, const size_type num_x, const size_type num_y,
                 const size_type num_z)
    : proc_info(proc_info), sim_pars(sim_pars), IM(num_x), JM(num_y), KM(num_z) {
  allocate_mem();
}

Block3d::~Block3d() { free_mem(); }

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
void Block3d::initial_condition() { status = cbns_init_tgv(ctx); }
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

// src/Block3d.cpp:318-357, one file per block.  mu is not in the file and is
// zero afterwards, as in the reference's restart path.
int Block3d::read_bin(const std::string fname) {
  std::vector<double> all((size_t)NEQ * IM * JM * KM);
  double t = 0.0;
  const int nb = num_blocks();
  for (int b = 0; b < nb; b++) {
    std::string fn;
    if (!block_file(fname, proc_info->rank, nb, b, &fn)) {
      std::cout << "checkpoint name " << fname << " lacks the rank suffix for " << nb << " blocks\n";
      return -1;
    }
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
  return status == CBNS_OK ? 0 : -1;
}

// src/Block3d.cpp:964-987, one file per block (ghosts included).
int Block3d::output_bin(const std::string fname) {
  const int nb = num_blocks();
  for (int b = 0; b < nb; b++) {
    std::string fn;
    if (!block_file(fname, proc_info->rank, nb, b, &fn)) {
      std::cout << "checkpoint name " << fname << " lacks the rank suffix for " << nb << " blocks\n";
      return -1;
    }
    int n[3], o[3];
    cbns_block_info(ctx, b, n, o);
    size_type d[3] = {(size_type)(n[0] + 2 * NG), (size_type)(n[1] + 2 * NG), (size_type)(n[2] + 2 * NG)};
    std::vector<double> Qg((size_t)NEQ * d[0] * d[1] * d[2]);
    if (cbns_get_block_state(ctx, b, CBNS_Q, Qg.data()) != CBNS_OK) return -1;
    std::ofstream fh(fn, std::ios::out | std::ios::binary);
    if (!fh) return -1;
    fh.write((const char*)&sim_pars->t_cur, sizeof(value_type));
    fh.write((const char*)d, sizeof d);
    fh.write((const char*)Qg.data(), Qg.size() * sizeof(double));
    if (!fh) return -1;
  }
  return 0;
}

void boundary_condition(Block3d* blk, value_type* Q) {
  blk->status = cbns_boundary_condition(blk->ctx, blk->which(Q));
}

void Block3d::calc_primitive(const value_type* Q) { status = cbns_calc_primitive(ctx, which(Q)); }
void Block3d::init_Q_p() { status = cbns_init_Q_p(ctx); }

// src/Block3d.cpp:396-469 for this rank's block
int Block3d::output_vtr(const std::string fname) { return output_vtr(fname, proc_info->rank); }

// src/Block3d.cpp:396-469 for block b
int Block3d::output_vtr(const std::string fname, int b) {
  int n[3], o[3];
  if (cbns_block_info(ctx, b, n, o) != CBNS_OK) return -1;
  const size_t G0 = n[0] + 2 * NG, G1 = n[1] + 2 * NG, G2 = n[2] + 2 * NG;
  std::vector<double> f[5];
  for (int a = 0; a < 5; a++) {
    f[a].resize(G0 * G1 * G2);
    if (cbns_get_block_primitive(ctx, b, a, f[a].data()) != CBNS_OK) return -1;
  }
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
  const std::vector<value_type>* xyz[3] = {&x, &y, &z};
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
  std::fclose(fh);
  return 0;
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

void Block3d::update_rk3(const value_type dt, const size_type stage) { status = cbns_update_rk3(ctx, dt, stage); }

int Block3d::step(value_type* dt_out) {
  cbns_set_time(ctx, sim_pars->t_cur);
  int fin = 0;
  status = cbns_step(ctx, dt_out, &fin);
  cbns_get_time(ctx, &sim_pars->t_cur);
  return status;
}
