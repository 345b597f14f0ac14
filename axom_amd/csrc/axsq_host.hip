// axsq_host.hip -- host-side parts of the boundary: the STL reader, the BVH's
// VTK output, and the quest signed-distance interface (the process-wide
// query behind signed_distance_* and its C names QUEST_signed_distance_*).
// No kernels live here; the device work goes through the entry points of
// axsq_api.hip.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "axsq_internal.hpp"
#include "axsq_quest.h"

namespace axsq {

int bvh_export_launch(const axsq_bvh* b, double* inner_nodes, int32_t* children,
                      int32_t* leaf_nodes, uint32_t* mcodes, hipStream_t s);

namespace {

// ---- STL (quest/readers/STLReader.cpp) ---------------------------------------
constexpr long kStlHeader = 80;  // binary: 80-byte header, int32 count,
constexpr long kStlTri = 50;     // then 50 bytes per triangle

// STLReader::isAsciiFormat (:40-86): ASCII unless the size matches the
// binary layout for the count at byte 80.
bool stl_is_binary(FILE* f, long size) {
  if (size < kStlHeader + 4) return false;
  int32_t ntris = 0;
  if (fseek(f, kStlHeader, SEEK_SET) != 0 || fread(&ntris, 4, 1, f) != 1) return false;
  return (int64_t)size == (int64_t)kStlHeader + 4 + (int64_t)ntris * kStlTri;
}

// A whole token that std::istream's double extraction reads completely:
// [+-] digits [. digits] [e|E [+-] digits], or [+-] . digits [exponent].
bool decimal_token(const std::string& t) {
  size_t i = 0, n = t.size();
  if (i < n && (t[i] == '+' || t[i] == '-')) ++i;
  size_t digits = 0;
  while (i < n && t[i] >= '0' && t[i] <= '9') ++i, ++digits;
  if (i < n && t[i] == '.') {
    ++i;
    while (i < n && t[i] >= '0' && t[i] <= '9') ++i, ++digits;
  }
  if (digits == 0) return false;
  if (i < n && (t[i] == 'e' || t[i] == 'E')) {
    ++i;
    if (i < n && (t[i] == '+' || t[i] == '-')) ++i;
    size_t ed = 0;
    while (i < n && t[i] >= '0' && t[i] <= '9') ++i, ++ed;
    if (ed == 0) return false;
  }
  return i == n;
}

// STLReader::readAsciiSTL (:89-124): the three tokens after each "vertex"
// token, as doubles. Reading stops at the first coordinate that is not a
// number (the reference's stream extraction fails there).
int stl_read_ascii(const char* file, std::vector<double>& nodes) {
  std::ifstream in(file);
  if (!in.is_open()) return AXSQ_ERR_INVALID_ARG;
  std::string tok;
  while (in >> tok) {
    if (tok != "vertex") continue;
    std::string c[3];
    if (!(in >> c[0] >> c[1] >> c[2])) break;
    if (!decimal_token(c[0]) || !decimal_token(c[1]) || !decimal_token(c[2])) break;
    for (int k = 0; k < 3; ++k) nodes.push_back(strtod(c[k].c_str(), nullptr));
  }
  return AXSQ_OK;
}

// STLReader::readBinarySTL (:127-190): float coordinates widened to double;
// normals and attributes skipped. The host is little-endian, as the format.
int stl_read_binary(FILE* f, std::vector<double>& nodes) {
  int32_t ntris = 0;
  if (fseek(f, kStlHeader, SEEK_SET) != 0 || fread(&ntris, 4, 1, f) != 1)
    return AXSQ_ERR_INVALID_ARG;
  nodes.reserve((size_t)ntris * 9);
  unsigned char rec[kStlTri];
  for (int32_t t = 0; t < ntris; ++t) {
    if (fread(rec, kStlTri, 1, f) != 1) return AXSQ_ERR_INVALID_ARG;
    for (int j = 0; j < 9; ++j) {
      float v;
      memcpy(&v, rec + 12 + 4 * j, 4);
      nodes.push_back(static_cast<double>(v));
    }
  }
  return AXSQ_OK;
}

int stl_read(const char* file, std::vector<double>& nodes) {
  nodes.clear();
  if (!file || !file[0]) {
    set_error("STL reader: empty file name");
    return AXSQ_ERR_INVALID_ARG;
  }
  FILE* f = fopen(file, "rb");
  if (!f) {
    set_error("STL reader: cannot open the file [%s]", file);
    return AXSQ_ERR_INVALID_ARG;
  }
  fseek(f, 0, SEEK_END);
  const long size = ftell(f);
  int rc;
  if (stl_is_binary(f, size)) {
    rc = stl_read_binary(f, nodes);
    fclose(f);
  } else {
    fclose(f);
    rc = stl_read_ascii(file, nodes);
  }
  if (rc != AXSQ_OK) set_error("STL reader: cannot read the file [%s]", file);
  return rc;
}

// Device copies of a triangle soup: SoA vertices and the implicit
// connectivity (3i, 3i+1, 3i+2) of STLReader::getMesh (:207-254).
struct DeviceSoup {
  double *x = nullptr, *y = nullptr, *z = nullptr;
  int32_t* conn = nullptr;
  int64_t nnodes = 0, nfaces = 0;
  ~DeviceSoup() {
    (void)hipFree(x);
    (void)hipFree(y);
    (void)hipFree(z);
    (void)hipFree(conn);
  }
  int upload(const std::vector<double>& nodes, hipStream_t s) {
    nnodes = (int64_t)nodes.size() / 3;
    nfaces = nnodes / 3;
    if (nfaces < 1) {
      set_error("STL mesh has no triangle");
      return AXSQ_ERR_INVALID_ARG;
    }
    std::vector<double> hx(nnodes), hy(nnodes), hz(nnodes);
    for (int64_t i = 0; i < nnodes; ++i) {
      hx[i] = nodes[3 * i];
      hy[i] = nodes[3 * i + 1];
      hz[i] = nodes[3 * i + 2];
    }
    std::vector<int32_t> hc(3 * nfaces);
    for (int64_t i = 0; i < 3 * nfaces; ++i) hc[i] = (int32_t)i;
    const size_t nb = sizeof(double) * nnodes;
    AXSQ_CHECK_HIP(hipMalloc((void**)&x, nb));
    AXSQ_CHECK_HIP(hipMalloc((void**)&y, nb));
    AXSQ_CHECK_HIP(hipMalloc((void**)&z, nb));
    AXSQ_CHECK_HIP(hipMalloc((void**)&conn, sizeof(int32_t) * hc.size()));
    AXSQ_CHECK_HIP(hipMemcpyAsync(x, hx.data(), nb, hipMemcpyHostToDevice, s));
    AXSQ_CHECK_HIP(hipMemcpyAsync(y, hy.data(), nb, hipMemcpyHostToDevice, s));
    AXSQ_CHECK_HIP(hipMemcpyAsync(z, hz.data(), nb, hipMemcpyHostToDevice, s));
    AXSQ_CHECK_HIP(hipMemcpyAsync(conn, hc.data(), sizeof(int32_t) * hc.size(),
                                  hipMemcpyHostToDevice, s));
    AXSQ_CHECK_HIP(hipStreamSynchronize(s));
    return AXSQ_OK;
  }
};

// ---- VTK output of a tree (spin/internal/linear_bvh/bvh_vtkio.hpp) -----------
struct VtkOut {
  std::ostringstream nodes, cells, levels;
  int32_t npoints = 0, nbins = 0;
  int dims = 3;
  void box(const double* lo, const double* hi) {  // lo, hi: dims entries each
    const int nn = (dims == 2) ? 4 : 8;
    if (dims == 2) {
      nodes << lo[0] << " " << lo[1] << " 0.0\n";
      nodes << hi[0] << " " << lo[1] << " 0.0\n";
      nodes << hi[0] << " " << hi[1] << " 0.0\n";
      nodes << lo[0] << " " << hi[1] << " 0.0\n";
    } else {
      for (int zz = 0; zz < 2; ++zz) {
        const double zc = zz ? hi[2] : lo[2];
        nodes << lo[0] << " " << lo[1] << " " << zc << "\n";
        nodes << hi[0] << " " << lo[1] << " " << zc << "\n";
        nodes << hi[0] << " " << hi[1] << " " << zc << "\n";
        nodes << lo[0] << " " << hi[1] << " " << zc << "\n";
      }
    }
    cells << nn << " ";
    for (int k = 0; k < nn; ++k) cells << (npoints + k) << " ";
    cells << "\n";
    ++nbins;
    npoints += nn;
  }
};

}  // namespace

int bvh_write_vtk(const axsq_bvh* b, const char* file) {
  const int64_t ninner = b->n - 1;
  const int bw = 2 * b->dims;  // doubles per exported box
  std::vector<double> boxes((size_t)(2 * ninner * bw));
  std::vector<int32_t> children((size_t)(2 * ninner));
  {
    double* dbox = nullptr;
    int32_t* dch = nullptr;
    AXSQ_CHECK_HIP(hipMalloc((void**)&dbox, sizeof(double) * boxes.size()));
    if (hipMalloc((void**)&dch, sizeof(int32_t) * children.size()) != hipSuccess) {
      (void)hipFree(dbox);
      set_error("axsq_bvh_write_vtk: out of device memory");
      return AXSQ_ERR_NOMEM;
    }
    int rc = bvh_export_launch(b, dbox, dch, nullptr, nullptr, nullptr);
    if (rc == AXSQ_OK &&
        (hipMemcpy(boxes.data(), dbox, sizeof(double) * boxes.size(), hipMemcpyDeviceToHost) !=
             hipSuccess ||
         hipMemcpy(children.data(), dch, sizeof(int32_t) * children.size(),
                   hipMemcpyDeviceToHost) != hipSuccess)) {
      set_error("axsq_bvh_write_vtk: copy of the tree failed");
      rc = AXSQ_ERR_HIP;
    }
    (void)hipFree(dbox);
    (void)hipFree(dch);
    if (rc != AXSQ_OK) return rc;
  }
  VtkOut v;
  v.dims = b->dims;
  // the root box at level 0 (bvh_vtkio.hpp write_root)
  const double root_lo[3] = {b->bounds[0], b->bounds[1], b->bounds[2]};
  const double root_hi[3] = {b->bounds[3], b->bounds[4], b->bounds[5]};
  v.box(root_lo, root_hi);
  v.levels << "0\n";
  // then each inner node's two child boxes, depth first, left subtree first
  // (write_recursive); `node` indexes the pair 2i of inner node i
  struct Item {
    int32_t node, level;
  };
  std::vector<Item> todo;
  if (ninner > 0) todo.push_back({0, 1});
  while (!todo.empty()) {
    const Item it = todo.back();
    todo.pop_back();
    for (int side = 0; side < 2; ++side) {
      const double* bx = boxes.data() + (size_t)(it.node + side) * bw;
      v.box(bx, bx + b->dims);
      v.levels << it.level << "\n";
    }
    const int32_t lc = children[(size_t)it.node], rc = children[(size_t)it.node + 1];
    if (rc > -1) todo.push_back({rc, it.level + 1});  // popped after the left subtree
    if (lc > -1) todo.push_back({lc, it.level + 1});
  }
  std::ofstream out(file);
  if (!out.is_open()) {
    set_error("axsq_bvh_write_vtk: cannot open [%s] for writing", file);
    return AXSQ_ERR_INVALID_ARG;
  }
  out << "# vtk DataFile Version 3.0\n";
  out << " BVHTree \n";
  out << "ASCII\n";
  out << "DATASET UNSTRUCTURED_GRID\n";
  out << "POINTS " << v.npoints << " double\n";
  out << v.nodes.str() << "\n";
  const int nn = (b->dims == 2) ? 4 : 8;
  out << "CELLS " << v.nbins << " " << v.nbins * (nn + 1) << "\n";
  out << v.cells.str() << "\n";
  out << "CELL_TYPES " << v.nbins << "\n";
  const int ctype = (b->dims == 2) ? 9 : 12;  // VTK_QUAD, VTK_HEXAHEDRON
  for (int32_t i = 0; i < v.nbins; ++i) out << ctype << "\n";
  out << "CELL_DATA " << v.nbins << "\n";
  out << "SCALARS level int\n";
  out << "LOOKUP_TABLE default\n";
  out << v.levels.str() << "\n";
  out << "\n";
  out.close();
  if (!out) {
    set_error("axsq_bvh_write_vtk: write to [%s] failed", file);
    return AXSQ_ERR_INVALID_ARG;
  }
  return AXSQ_OK;
}

// ---- the signed-distance interface (quest/interface/signed_distance.cpp) ------
namespace {

// Options (signed_distance.cpp:84-100 parameters_t, with its defaults) and
// the query.
struct SdParams {
  int dimension = 3;
  bool verbose = false;
  bool is_closed_surface = true;
  bool use_shared_memory = false;
  bool compute_sign = true;
  int allocator_id = -1;
  int exec_space = QUEST_SignedDistExec_CPU;
  int fp_mode = AXSQ_FP64;
};
std::mutex g_sd_mu;
SdParams g_sd;
axsq_sdf* g_sd_query = nullptr;

void sd_warn(const char* msg) {
  if (g_sd.verbose) fprintf(stderr, "[quest signed distance] %s\n", msg);
}

// SLIC_ERROR_IF(signed_distance_initialized(), ...) of every option setter
bool sd_settable(const char* name) {
  if (g_sd_query) {
    set_error("%s: signed distance query already initialized; setting option has no effect",
              name);
    return false;
  }
  return true;
}

bool sd_ready(const char* name) {
  if (!g_sd_query) {
    set_error("%s: signed distance query must be initialized first", name);
    return false;
  }
  return true;
}

// Device copy of n values at p when p is host memory (tmp owns it); p itself
// when the device can read it.
template <class T>
int device_view(const T* p, int64_t n, T** tmp, const T** out) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) == hipSuccess &&
      (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged)) {
    *out = p;
    return AXSQ_OK;
  }
  (void)hipGetLastError();  // an unregistered host pointer is not an error here
  AXSQ_CHECK_HIP(hipMalloc((void**)tmp, sizeof(T) * (size_t)(n > 0 ? n : 1)));
  AXSQ_CHECK_HIP(hipMemcpy(*tmp, p, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  *out = *tmp;
  return AXSQ_OK;
}

int sd_init_mesh(const double* x, const double* y, const double* z, int64_t nnodes,
                 const int32_t* conn, int64_t ncells) {
  if (g_sd_query) {
    set_error("signed_distance_init: signed distance query has already been initialized");
    return -1;
  }
  if (!x || !y || !z || !conn || nnodes < 3 || ncells < 1) {
    set_error("signed_distance_init: need a triangle mesh with >= 3 nodes and >= 1 cell");
    return -1;
  }
  double *tx = nullptr, *ty = nullptr, *tz = nullptr;
  int32_t* tc = nullptr;
  const double *dx, *dy, *dz;
  const int32_t* dc;
  int rc = device_view(x, nnodes, &tx, &dx);
  if (rc == AXSQ_OK) rc = device_view(y, nnodes, &ty, &dy);
  if (rc == AXSQ_OK) rc = device_view(z, nnodes, &tz, &dz);
  if (rc == AXSQ_OK) rc = device_view(conn, 3 * ncells, &tc, &dc);
  if (rc == AXSQ_OK)
    rc = axsq_sdf_create(&g_sd_query, dx, dy, dz, nnodes, dc, ncells,
                         g_sd.is_closed_surface ? 1 : 0, g_sd.compute_sign ? 1 : 0, g_sd.fp_mode,
                         nullptr);
  (void)hipFree(tx);
  (void)hipFree(ty);
  (void)hipFree(tz);
  (void)hipFree(tc);
  if (rc != AXSQ_OK) {
    g_sd_query = nullptr;
    sd_warn(axsq_last_error());
    return -1;
  }
  return 0;
}

}  // namespace
}  // namespace axsq

using namespace axsq;

extern "C" {

int axsq_host_free(void* p) {
  free(p);
  return AXSQ_OK;
}

int axsq_stl_read(const char* file, double** nodes, int64_t* nnodes, int64_t* nfaces) {
  clear_error();
  if (!nodes || !nnodes || !nfaces) {
    set_error("axsq_stl_read: NULL output");
    return AXSQ_ERR_INVALID_ARG;
  }
  *nodes = nullptr;
  *nnodes = *nfaces = 0;
  std::vector<double> v;
  AXSQ_TRY(stl_read(file, v));
  double* out = (double*)malloc(sizeof(double) * (v.empty() ? 1 : v.size()));
  if (!out) {
    set_error("axsq_stl_read: out of host memory");
    return AXSQ_ERR_NOMEM;
  }
  if (!v.empty()) memcpy(out, v.data(), sizeof(double) * v.size());
  *nodes = out;
  *nnodes = (int64_t)v.size() / 3;
  *nfaces = *nnodes / 3;
  return AXSQ_OK;
}

int axsq_sdf_create_from_stl(axsq_sdf** out, const char* file, int watertight,
                             int compute_sign, int fp_mode, void* stream) {
  clear_error();
  if (!out) {
    set_error("axsq_sdf_create_from_stl: out is NULL");
    return AXSQ_ERR_INVALID_ARG;
  }
  *out = nullptr;
  std::vector<double> v;
  AXSQ_TRY(stl_read(file, v));
  DeviceSoup m;
  AXSQ_TRY(m.upload(v, (hipStream_t)stream));
  return axsq_sdf_create(out, m.x, m.y, m.z, m.nnodes, m.conn, m.nfaces, watertight,
                         compute_sign, fp_mode, stream);
}

int axsq_bvh_write_vtk(const axsq_bvh* b, const char* file) {
  clear_error();
  if (!b || !b->nodes || !file) {
    set_error("axsq_bvh_write_vtk: tree not built or no file name");
    return AXSQ_ERR_INVALID_ARG;
  }
  return bvh_write_vtk(b, file);
}

// -- QUEST_signed_distance_* (quest/interface/c_fortran/wrapQUEST.h:84-129 over
// quest/interface/signed_distance.cpp) --------------------------------------------

int QUEST_signed_distance_init_serial(const char* file) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (g_sd_query) {
    set_error("signed_distance_init: signed distance query has already been initialized");
    return -1;
  }
  if (g_sd.dimension != 3) {  // signed_distance.cpp:113-117
    set_error("signed_distance_init: the SignedDistance query is only supported in 3D");
    sd_warn(axsq_last_error());
    return -1;
  }
  std::vector<double> v;
  DeviceSoup m;
  if (stl_read(file, v) != AXSQ_OK || m.upload(v, nullptr) != AXSQ_OK) {
    sd_warn(axsq_last_error());
    return -1;
  }
  return sd_init_mesh(m.x, m.y, m.z, m.nnodes, m.conn, m.nfaces);
}

int QUEST_signed_distance_init_serial_bufferify(char* file, int file_len) {
  // a Fortran string: its length without trailing blanks
  int n = file ? file_len : 0;
  while (n > 0 && file[n - 1] == ' ') --n;
  const std::string f(file ? file : "", (size_t)(n > 0 ? n : 0));
  return QUEST_signed_distance_init_serial(f.c_str());
}

int axsq_signed_distance_init_mesh(const double* x, const double* y, const double* z,
                                   int64_t nnodes, const int32_t* conn, int64_t ncells) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (g_sd.dimension != 3) {
    set_error("signed_distance_init: the SignedDistance query is only supported in 3D");
    return -1;
  }
  return sd_init_mesh(x, y, z, nnodes, conn, ncells);
}

bool QUEST_signed_distance_initialized(void) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  return g_sd_query != nullptr;
}

void QUEST_signed_distance_get_mesh_bounds(double* lo, double* hi) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_ready("signed_distance_get_mesh_bounds")) return;
  if (!lo || !hi) {
    set_error("signed_distance_get_mesh_bounds: supplied buffer is null");
    return;
  }
  (void)axsq_sdf_mesh_bounds(g_sd_query, lo, hi);
}

void QUEST_signed_distance_set_dimension(int dim) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (dim != 3) {
    set_error("signed_distance_set_dimension: the signed distance query only supports 3D");
    return;
  }
  if (sd_settable("signed_distance_set_dimension")) g_sd.dimension = dim;
}

void QUEST_signed_distance_set_closed_surface(bool status) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (sd_settable("signed_distance_set_closed_surface")) g_sd.is_closed_surface = status;
}

void QUEST_signed_distance_set_compute_signs(bool computeSign) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (sd_settable("signed_distance_set_compute_signs")) g_sd.compute_sign = computeSign;
}

void QUEST_signed_distance_set_allocator(int allocatorID) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (sd_settable("signed_distance_set_allocator")) g_sd.allocator_id = allocatorID;
}

void QUEST_signed_distance_set_verbose(bool status) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (sd_settable("signed_distance_set_verbose")) g_sd.verbose = status;
}

void QUEST_signed_distance_use_shared_memory(bool status) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_settable("signed_distance_use_shared_memory")) return;
  g_sd.use_shared_memory = status;
  // MPI-3 shared memory is not part of this build (signed_distance.cpp:289-291)
  if (status) sd_warn("shared memory requires MPI-3; the option is ignored");
}

void QUEST_signed_distance_set_execution_space(int execSpace) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_settable("signed_distance_set_execution_space")) return;
  if (execSpace != QUEST_SignedDistExec_CPU && execSpace != QUEST_SignedDistExec_OpenMP &&
      execSpace != QUEST_SignedDistExec_GPU) {
    set_error("signed_distance_set_execution_space: unknown execution space %d", execSpace);
    return;
  }
  g_sd.exec_space = execSpace;
}

void axsq_signed_distance_set_precision(int fp_mode) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_settable("signed_distance_set_precision")) return;
  if (fp_mode != AXSQ_FP64 && fp_mode != AXSQ_FP32_TRAVERSAL) {
    set_error("signed_distance_set_precision: unknown fp_mode %d", fp_mode);
    return;
  }
  g_sd.fp_mode = fp_mode;
}

double QUEST_signed_distance_evaluate_0(double x, double y, double z) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_ready("signed_distance_evaluate")) return NAN;
  const double q[3] = {x, y, z};
  double d = NAN;
  if (axsq_sdf_eval_point(g_sd_query, q, g_sd.fp_mode, &d, nullptr, nullptr) != AXSQ_OK)
    return NAN;
  return d;
}

double QUEST_signed_distance_evaluate_1(double x, double y, double z, double* cp_x,
                                        double* cp_y, double* cp_z, double* n_x, double* n_y,
                                        double* n_z) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_ready("signed_distance_evaluate")) return NAN;
  if (!cp_x || !cp_y || !cp_z || !n_x || !n_y || !n_z) {
    set_error("signed_distance_evaluate: NULL output");
    return NAN;
  }
  const double q[3] = {x, y, z};
  double d = NAN, cp[3], n[3];
  if (axsq_sdf_eval_point(g_sd_query, q, g_sd.fp_mode, &d, cp, n) != AXSQ_OK) return NAN;
  *cp_x = cp[0];
  *cp_y = cp[1];
  *cp_z = cp[2];
  *n_x = n[0];
  *n_y = n[1];
  *n_z = n[2];
  return d;
}

int axsq_signed_distance_evaluate_batch(const double* x, const double* y, const double* z,
                                        int npoints, double* phi) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  if (!sd_ready("signed_distance_evaluate")) return AXSQ_ERR_STATE;
  if (!x || !y || !z || !phi || npoints < 0) {
    set_error("signed_distance_evaluate: a coordinate or output array is null");
    return AXSQ_ERR_INVALID_ARG;
  }
  if (npoints == 0) return AXSQ_OK;
  double *tx = nullptr, *ty = nullptr, *tz = nullptr, *tphi = nullptr;
  const double *dx, *dy, *dz, *dphi;
  int rc = device_view(x, npoints, &tx, &dx);
  if (rc == AXSQ_OK) rc = device_view(y, npoints, &ty, &dy);
  if (rc == AXSQ_OK) rc = device_view(z, npoints, &tz, &dz);
  if (rc == AXSQ_OK) rc = device_view((const double*)phi, npoints, &tphi, &dphi);
  if (rc == AXSQ_OK)
    rc = axsq_sdf_eval(g_sd_query, dx, dy, dz, 1, npoints, g_sd.fp_mode, (double*)dphi,
                       nullptr, nullptr, nullptr, nullptr);
  if (rc == AXSQ_OK && tphi &&
      hipMemcpy(phi, tphi, sizeof(double) * npoints, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("signed_distance_evaluate: copy of the distances failed");
    rc = AXSQ_ERR_HIP;
  }
  (void)hipFree(tx);
  (void)hipFree(ty);
  (void)hipFree(tz);
  (void)hipFree(tphi);
  return rc;
}

void QUEST_signed_distance_finalize(void) {
  std::lock_guard<std::mutex> g(g_sd_mu);
  clear_error();
  (void)axsq_sdf_destroy(g_sd_query);
  g_sd_query = nullptr;
}

}  // extern "C"
