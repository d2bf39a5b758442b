// Host-side mesh and operator handling for libdgk.
//
//  * load_operators      : operator_data/p{p}.dat reader (element_init,
//                          mod_mesh.f90:253-263).
//  * dgk_mesh_build      : the reference's per-node mesh arrays for a periodic
//                          box (mesh_setup, mod_mesh.f90:74-193), for callers
//                          that do not have the reference host at hand.
//  * build_tables_*      : compression of those arrays into what the kernels
//                          read: a 6-entry neighbour table per element (or a
//                          resolved per-face-node index), the halo folded into
//                          it, and deduplicated per-element geometry records.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

// One list-directed record: the first n numbers of the next line(s).
bool read_record(std::ifstream &in, double *dst, size_t n) {
  size_t got = 0;
  std::string line;
  while (got < n) {
    if (!std::getline(in, line)) return false;
    const char *s = line.c_str();
    char *end = nullptr;
    while (got < n) {
      double v = std::strtod(s, &end);
      if (end == s) break;
      dst[got++] = v;
      s = end;
    }
  }
  return true;
}

// Reference face order -y, +x, +y, -x, -z, +z: normal sign (axis: kFaceAxis).
constexpr double kFaceSign[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};

// Periodic neighbour of element (i,j,k) across face f (0-based indices,
// eval_neighbor_element, mod_mesh.f90:407-483).  Returns the neighbour's
// element index and whether the face lies on the domain boundary.
int64_t neighbour(int i, int j, int k, int f, int NeX, int NeY, int NeZ,
                  bool &boundary) {
  boundary = false;
  switch (f) {
  case 0: if (j > 0) --j; else { j = NeY - 1; boundary = true; } break;
  case 1: if (i < NeX - 1) ++i; else { i = 0; boundary = true; } break;
  case 2: if (j < NeY - 1) ++j; else { j = 0; boundary = true; } break;
  case 3: if (i > 0) --i; else { i = NeX - 1; boundary = true; } break;
  case 4: if (k > 0) --k; else { k = NeZ - 1; boundary = true; } break;
  default: if (k < NeZ - 1) ++k; else { k = 0; boundary = true; } break;
  }
  return i + (int64_t)j * NeX + (int64_t)k * NeX * NeY;
}

struct RecHash {
  size_t operator()(const std::array<uint64_t, 27> &a) const {
    uint64_t h = 1469598103934665603ULL;
    for (uint64_t v : a) { h ^= v; h *= 1099511628211ULL; }
    return (size_t)h;
  }
};

}  // namespace

void fill_box_geo(double dx, double dy, double dz, double *rec) {
  const double h[3] = {dx, dy, dz};
  std::memset(rec, 0, sizeof(double) * kGeoStride);
  for (int d = 0; d < 3; ++d) rec[kGeoEscale + d] = 2.0 / h[d];
  for (int f = 0; f < 6; ++f) {
    rec[kGeoFscale + f] = 2.0 / h[kFaceAxis[f]];
    rec[kGeoNormal + 3 * f + kFaceAxis[f]] = kFaceSign[f];
  }
}

bool geo_axis_aligned(const std::vector<double> &geo) {
  for (size_t r = 0; r + kGeoStride <= geo.size(); r += kGeoStride)
    for (int f = 0; f < 6; ++f)
      for (int d = 0; d < 3; ++d)
        if (d != kFaceAxis[f] && geo[r + kGeoNormal + 3 * f + d] != 0.0) return false;
  return true;
}

std::vector<int32_t> brick_order(int NeX, int NeY, int NeZ, int bx, int by, int bz) {
  bx = std::max(1, std::min(bx, NeX));
  by = std::max(1, std::min(by, NeY));
  bz = std::max(1, std::min(bz, NeZ));
  std::vector<int32_t> order;
  order.reserve((size_t)NeX * NeY * NeZ);
  for (int K = 0; K < NeZ; K += bz)
    for (int J = 0; J < NeY; J += by)
      for (int I = 0; I < NeX; I += bx)
        for (int k = K; k < std::min(K + bz, NeZ); ++k)
          for (int j = J; j < std::min(J + by, NeY); ++j)
            for (int i = I; i < std::min(I + bx, NeX); ++i)
              order.push_back((int32_t)(i + (int64_t)j * NeX + (int64_t)k * NeX * NeY));
  return order;
}

int64_t z_plane_size(const MeshTables &t) {
  if (t.nbr.empty() || t.Ne < 2) return 0;
  const int64_t P = t.nbr[5];
  if (P <= 0 || P >= t.Ne || t.Ne % P != 0) return 0;
  for (int64_t e = 0; e < t.Ne; ++e)
    if (t.nbr[(size_t)e * 6 + 5] != (e + P) % t.Ne) return 0;
  return P;
}

int64_t xcd_period(int64_t P) { return P >= 8 ? P : 0; }

void default_order(MeshTables &t) {
  t.order.clear();
  t.zperiod = 0;
  const char *env = std::getenv("DGK_ORDER");
  if (env && std::string(env) == "identity") return;
  const int64_t P = z_plane_size(t);
  if (P == 0) return;
  long q = 0;
  if (env && std::sscanf(env, "zperiod:%ld", &q) == 1) {  // a period of choice
    if (q >= 8 && q <= t.Ne && t.Ne % q == 0) t.zperiod = q;
    return;
  }
  t.zperiod = xcd_period(P);  // the kernels compute this order; no array
}

int load_operators(const char *dir, int p, std::vector<double> &x1D,
                   std::vector<double> &D, std::vector<double> &Lift) {
  if (!dir) return set_error(DGK_ERR_ARG, "operator directory is NULL");
  if (p < kMinNq - 1 || p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order " + std::to_string(p) +
                                              " outside 1.." + std::to_string(kMaxNq - 1));
  const int Nq = p + 1;
  const std::string path = std::string(dir) + "/p" + std::to_string(p) + ".dat";
  std::ifstream in(path);
  // orders beyond the stored files use the closed forms
  if (!in && Nq > kMaxNqFile) return make_operators(p, x1D, D, Lift);
  if (!in) return set_error(DGK_ERR_IO, "cannot open operator file " + path);
  x1D.assign(Nq, 0.0);
  D.assign((size_t)Nq * Nq, 0.0);
  Lift.assign((size_t)Nq * Nq * Nq * 6, 0.0);
  if (!read_record(in, x1D.data(), x1D.size()) ||
      !read_record(in, D.data(), D.size()) ||
      !read_record(in, Lift.data(), Lift.size()))
    return set_error(DGK_ERR_IO, "operator file " + path + " is truncated");
  return DGK_OK;
}

int build_tables_structured(int NeX, int NeY, int NeZ, int p,
                            const double *D1D, const double *Lift_mat,
                            double dom_x, double dom_y, double dom_z,
                            MeshTables &t) {
  if (NeX < 1 || NeY < 1 || NeZ < 1)
    return set_error(DGK_ERR_ARG, "element counts must be >= 1");
  if (p < kMinNq - 1 || p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order outside 1.." + std::to_string(kMaxNq - 1));
  if (!(dom_x > 0 && dom_y > 0 && dom_z > 0))
    return set_error(DGK_ERR_ARG, "domain sizes must be positive");
  const int Nq = p + 1;
  t.Nq = Nq; t.Np = Nq * Nq * Nq; t.Nfp = Nq * Nq; t.NfpTot = 6 * t.Nfp;
  t.Ne = (int64_t)NeX * NeY * NeZ;
  const int64_t Nhalo = 2 * ((int64_t)NeY * NeZ + (int64_t)NeX * NeZ + (int64_t)NeX * NeY);
  t.NeA = t.Ne + Nhalo;
  t.NhaloNode = (int64_t)t.Nfp * Nhalo;
  if ((int64_t)t.Np * t.NeA >= (int64_t)INT32_MAX)
    return set_error(DGK_ERR_UNSUPPORTED, "mesh exceeds 2^31 nodes (int32 maps)");
  t.D.assign(D1D, D1D + (size_t)Nq * Nq);
  t.Lift.assign(Lift_mat, Lift_mat + (size_t)t.Np * 6);
  t.nbr.resize((size_t)t.Ne * 6);
  t.vmapP.clear();
  t.halo_src.resize(t.NhaloNode);
  int64_t ibuf = 0;
  for (int k = 0; k < NeZ; ++k)
    for (int j = 0; j < NeY; ++j)
      for (int i = 0; i < NeX; ++i) {
        const int64_t e = i + (int64_t)j * NeX + (int64_t)k * NeX * NeY;
        for (int f = 0; f < 6; ++f) {
          bool bnd;
          const int64_t nb = neighbour(i, j, k, f, NeX, NeY, NeZ, bnd);
          t.nbr[e * 6 + f] = (int32_t)nb;
          if (bnd)  // halo slots in serialized (k,j,i,f,fp) order
            for (int fp = 0; fp < t.Nfp; ++fp)
              t.halo_src[ibuf++] =
                  (int32_t)(nb * t.Np + nbr_face_node(f, fp % Nq, fp / Nq, Nq));
        }
      }
  t.geo.assign(kGeoStride, 0.0);
  fill_box_geo(dom_x / (double)NeX, dom_y / (double)NeY, dom_z / (double)NeZ,
               t.geo.data());
  t.geoid.clear();
  int bx = 0, by = 0, bz = 0;
  const char *env = std::getenv("DGK_ORDER");
  if (env && std::sscanf(env, "brick:%d,%d,%d", &bx, &by, &bz) == 3 && bx > 0 && by > 0 &&
      bz > 0)
  {
    t.order = brick_order(NeX, NeY, NeZ, bx, by, bz);
    t.zperiod = 0;
  }
  else
    default_order(t);
  t.axis_aligned = geo_axis_aligned(t.geo);
  return DGK_OK;
}

int build_tables_batch(int p, const double *D1D, const double *Lift_mat,
                       const std::vector<BatchInstance> &inst, MeshTables &t,
                       std::vector<int64_t> &elem_offset) {
  if (p < kMinNq - 1 || p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order outside 1.." + std::to_string(kMaxNq - 1));
  if (inst.empty()) return set_error(DGK_ERR_ARG, "empty batch");
  const int Nq = p + 1;
  t.Nq = Nq; t.Np = Nq * Nq * Nq; t.Nfp = Nq * Nq; t.NfpTot = 6 * t.Nfp;
  int64_t total = 0;
  elem_offset.assign(inst.size() + 1, 0);
  for (size_t k = 0; k < inst.size(); ++k) {
    const BatchInstance &b = inst[k];
    if (b.NeX < 1 || b.NeY < 1 || b.NeZ < 1)
      return set_error(DGK_ERR_ARG, "instance " + std::to_string(k) + ": element counts must be >= 1");
    if (!(b.dom[0] > 0 && b.dom[1] > 0 && b.dom[2] > 0) || !(b.dt > 0))
      return set_error(DGK_ERR_ARG, "instance " + std::to_string(k) + ": domain and dt must be positive");
    elem_offset[k] = total;
    total += (int64_t)b.NeX * b.NeY * b.NeZ;
  }
  elem_offset[inst.size()] = total;
  if ((int64_t)t.Np * total >= (int64_t)INT32_MAX)
    return set_error(DGK_ERR_UNSUPPORTED, "batch exceeds 2^31 nodes (int32 maps)");
  t.Ne = total; t.NeA = total; t.NhaloNode = 0;
  t.D.assign(D1D, D1D + (size_t)Nq * Nq);
  t.Lift.assign(Lift_mat, Lift_mat + (size_t)t.Np * 6);
  t.nbr.resize((size_t)total * 6);
  t.vmapP.clear();
  t.halo_src.clear();
  t.geo.assign(inst.size() * kGeoStride, 0.0);
  t.geoid.resize(total);
  for (size_t k = 0; k < inst.size(); ++k) {
    const BatchInstance &b = inst[k];
    double *rec = &t.geo[k * kGeoStride];
    fill_box_geo(b.dom[0] / b.NeX, b.dom[1] / b.NeY, b.dom[2] / b.NeZ, rec);
    rec[kGeoDt] = b.dt;
    const int64_t off = elem_offset[k];
    for (int kk = 0; kk < b.NeZ; ++kk)
      for (int j = 0; j < b.NeY; ++j)
        for (int i = 0; i < b.NeX; ++i) {
          const int64_t e = i + (int64_t)j * b.NeX + (int64_t)kk * b.NeX * b.NeY;
          for (int f = 0; f < 6; ++f) {
            bool bnd;
            t.nbr[(off + e) * 6 + f] = (int32_t)(off + neighbour(i, j, kk, f, b.NeX, b.NeY, b.NeZ, bnd));
          }
          t.geoid[off + e] = (uint32_t)k;
        }
  }
  t.order.clear();
  t.zperiod = 0;
  t.axis_aligned = true;
  return DGK_OK;
}

int build_tables_from_reference(int Nq, int64_t Ne, int64_t NeA,
                                int64_t NhaloNode, const double *D1D,
                                const double *Lift_mat, const int *VMapM,
                                const int *VMapP, const int *halo_src_map,
                                const double *normal_fn, const double *Escale,
                                const double *Fscale, MeshTables &t) {
  if (!D1D || !Lift_mat || !VMapM || !VMapP || !normal_fn || !Escale || !Fscale ||
      (NhaloNode > 0 && !halo_src_map))
    return set_error(DGK_ERR_ARG, "null array argument");
  if (Nq < kMinNq || Nq > kMaxNq)
    return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2.." + std::to_string(kMaxNq));
  if (Ne < 1 || NeA < Ne || NhaloNode < 0)
    return set_error(DGK_ERR_ARG, "inconsistent Ne / NeA / NhaloNode");
  const int Np = Nq * Nq * Nq, Nfp = Nq * Nq, NfpTot = 6 * Nfp;
  if ((int64_t)Np * Ne + NhaloNode > (int64_t)Np * NeA)
    return set_error(DGK_ERR_ARG, "halo does not fit in Np*NeA");
  if ((int64_t)Np * NeA >= (int64_t)INT32_MAX)
    return set_error(DGK_ERR_UNSUPPORTED, "mesh exceeds 2^31 nodes (int32 maps)");
  t.Nq = Nq; t.Np = Np; t.Nfp = Nfp; t.NfpTot = NfpTot;
  t.Ne = Ne; t.NeA = NeA; t.NhaloNode = NhaloNode;
  t.D.assign(D1D, D1D + (size_t)Nq * Nq);
  t.Lift.assign(Lift_mat, Lift_mat + (size_t)Np * 6);

  const int64_t owned = (int64_t)Np * Ne;
  t.halo_src.resize(NhaloNode);
  for (int64_t ib = 0; ib < NhaloNode; ++ib) {
    const int64_t s = (int64_t)halo_src_map[ib] - 1;
    if (s < 0 || s >= owned)
      return set_error(DGK_ERR_UNSUPPORTED,
                       "halo_src_map entry " + std::to_string(ib + 1) +
                           " does not address an owned node");
    t.halo_src[ib] = (int32_t)s;
  }

  // VMapM must address the element's own face nodes in Fmask order
  // (mod_mesh.f90:159); VMapP is resolved through the halo (update_halo
  // followed by the read at VMapP equals a direct read at the source).
  std::vector<int32_t> resolved((size_t)NfpTot * Ne);
  bool compact = true;
  std::vector<int32_t> nbr((size_t)Ne * 6);
  for (int64_t e = 0; e < Ne; ++e) {
    for (int f = 0; f < 6; ++f) {
      int64_t nb = -1;
      for (int fp = 0; fp < Nfp; ++fp) {
        const int a = fp % Nq, b = fp / Nq;
        const int64_t r = (int64_t)f * Nfp + fp + e * NfpTot;
        const int64_t m = (int64_t)VMapM[r] - 1;
        if (m != e * Np + own_face_node(f, a, b, Nq))
          return set_error(DGK_ERR_UNSUPPORTED,
                           "VMapM does not address the element's own Fmask node "
                           "at face node " + std::to_string(r + 1));
        int64_t pidx = (int64_t)VMapP[r] - 1;
        if (pidx < 0 || pidx >= owned + NhaloNode)
          return set_error(DGK_ERR_ARG, "VMapP entry " + std::to_string(r + 1) +
                                            " out of range");
        if (pidx >= owned) pidx = t.halo_src[pidx - owned];
        resolved[r] = (int32_t)pidx;
        if (compact) {
          const int64_t off = pidx - nbr_face_node(f, a, b, Nq);
          if (off < 0 || off % Np != 0) compact = false;
          else if (fp == 0) nb = off / Np;
          else if (off / Np != nb) compact = false;
        }
      }
      if (compact) nbr[e * 6 + f] = (int32_t)nb;
    }
  }
  if (compact) { t.nbr.swap(nbr); t.vmapP.clear(); }
  else { t.vmapP.swap(resolved); t.nbr.clear(); }

  // Geometry: Escale constant per element and direction, Fscale and the
  // normal constant per face (affine hexahedra, mod_mesh.f90:147-149,364-394).
  std::unordered_map<std::array<uint64_t, 27>, uint32_t, RecHash> ids;
  std::vector<double> geo;
  std::vector<uint32_t> geoid(Ne);
  std::array<double, 27> rec;
  for (int64_t e = 0; e < Ne; ++e) {
    for (int d = 0; d < 3; ++d) {
      const double *es = Escale + e * Np + (int64_t)d * Np * Ne;
      for (int n = 1; n < Np; ++n)
        if (es[n] != es[0])
          return set_error(DGK_ERR_UNSUPPORTED,
                           "Escale varies inside element " + std::to_string(e + 1) +
                               " (only affine hexahedra are supported)");
      rec[kGeoEscale + d] = es[0];
    }
    for (int f = 0; f < 6; ++f) {
      const int64_t r0 = (int64_t)f * Nfp + e * NfpTot;
      for (int fp = 1; fp < Nfp; ++fp) {
        bool same = Fscale[r0 + fp] == Fscale[r0];
        for (int d = 0; d < 3; ++d)
          same = same && normal_fn[r0 + fp + (int64_t)d * NfpTot * Ne] ==
                             normal_fn[r0 + (int64_t)d * NfpTot * Ne];
        if (!same)
          return set_error(DGK_ERR_UNSUPPORTED,
                           "Fscale/normal_fn vary across face " + std::to_string(f + 1) +
                               " of element " + std::to_string(e + 1));
      }
      rec[kGeoFscale + f] = Fscale[r0];
      for (int d = 0; d < 3; ++d)
        rec[kGeoNormal + 3 * f + d] = normal_fn[r0 + (int64_t)d * NfpTot * Ne];
    }
    std::array<uint64_t, 27> key;
    std::memcpy(key.data(), rec.data(), sizeof(key));
    auto it = ids.find(key);
    uint32_t id;
    if (it == ids.end()) {
      id = (uint32_t)ids.size();
      ids.emplace(key, id);
      geo.resize((size_t)(id + 1) * kGeoStride, 0.0);
      std::memcpy(&geo[(size_t)id * kGeoStride], rec.data(), sizeof(rec));
    } else {
      id = it->second;
    }
    geoid[e] = id;
  }
  t.geo.swap(geo);
  if (ids.size() == 1) t.geoid.clear();
  else t.geoid.swap(geoid);
  t.axis_aligned = geo_axis_aligned(t.geo);
  default_order(t);
  return DGK_OK;
}

}  // namespace dgk

using namespace dgk;

extern "C" int dgk_load_operators(const char *dir, int p, double *x1D,
                                  double *D1D, double *Lift_mat) {
  clear_error();
  if (!x1D || !D1D || !Lift_mat) return set_error(DGK_ERR_ARG, "null output array");
  std::vector<double> x, D, L;
  DGK_TRY(load_operators(dir, p, x, D, L));
  std::memcpy(x1D, x.data(), x.size() * sizeof(double));
  std::memcpy(D1D, D.data(), D.size() * sizeof(double));
  std::memcpy(Lift_mat, L.data(), L.size() * sizeof(double));
  return DGK_OK;
}

extern "C" int dgk_element_sizes(int p, int *Nq, int *Np, int *Nfp, int *NfpTot) {
  clear_error();
  if (p < kMinNq - 1 || p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order outside 1.." + std::to_string(kMaxNq - 1));
  const int q = p + 1;
  if (Nq) *Nq = q;
  if (Np) *Np = q * q * q;
  if (Nfp) *Nfp = q * q;
  if (NfpTot) *NfpTot = 6 * q * q;
  return DGK_OK;
}

extern "C" int dgk_mesh_sizes(int NeX, int NeY, int NeZ, int p, int64_t *Ne,
                              int64_t *NeA, int64_t *NhaloNode) {
  clear_error();
  if (NeX < 1 || NeY < 1 || NeZ < 1) return set_error(DGK_ERR_ARG, "element counts must be >= 1");
  if (p < kMinNq - 1 || p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order outside 1.." + std::to_string(kMaxNq - 1));
  const int64_t ne = (int64_t)NeX * NeY * NeZ;
  const int64_t nh = 2 * ((int64_t)NeY * NeZ + (int64_t)NeX * NeZ + (int64_t)NeX * NeY);
  if (Ne) *Ne = ne;
  if (NeA) *NeA = ne + nh;
  if (NhaloNode) *NhaloNode = (int64_t)(p + 1) * (p + 1) * nh;
  return DGK_OK;
}

extern "C" int dgk_mesh_build(int NeX, int NeY, int NeZ, int p,
                              const double *x1D, double dom_x, double dom_y,
                              double dom_z, int *VMapM, int *VMapP,
                              int *halo_src_map, double *normal_fn,
                              double *Escale, double *Fscale, double *pos_en) {
  clear_error();
  if (!x1D || !VMapM || !VMapP || !halo_src_map || !normal_fn || !Escale || !Fscale)
    return set_error(DGK_ERR_ARG, "null array argument");
  int64_t Ne, NeA, NhaloNode;
  DGK_TRY(dgk_mesh_sizes(NeX, NeY, NeZ, p, &Ne, &NeA, &NhaloNode));
  const int Nq = p + 1, Np = Nq * Nq * Nq, Nfp = Nq * Nq, NfpTot = 6 * Nfp;
  if ((int64_t)Np * NeA >= (int64_t)INT32_MAX)
    return set_error(DGK_ERR_UNSUPPORTED, "mesh exceeds 2^31 nodes (int32 maps)");
  const double h[3] = {dom_x / (double)NeX, dom_y / (double)NeY, dom_z / (double)NeZ};
  double geo[kGeoStride];
  fill_box_geo(h[0], h[1], h[2], geo);
  int64_t ibuf = 0;
  for (int k = 0; k < NeZ; ++k)
    for (int j = 0; j < NeY; ++j)
      for (int i = 0; i < NeX; ++i) {
        const int64_t e = i + (int64_t)j * NeX + (int64_t)k * NeX * NeY;
        const double c[3] = {((double)(i + 1) - 0.5) * h[0], ((double)(j + 1) - 0.5) * h[1],
                             ((double)(k + 1) - 0.5) * h[2]};
        for (int n = 0; n < Np; ++n) {
          const int a[3] = {n % Nq, (n / Nq) % Nq, n / (Nq * Nq)};
          for (int d = 0; d < 3; ++d) {
            Escale[n + e * Np + (int64_t)d * Np * Ne] = geo[kGeoEscale + d];
            if (pos_en)
              pos_en[n + e * Np + (int64_t)d * Np * Ne] = c[d] + 0.5 * h[d] * x1D[a[d]];
          }
        }
        for (int f = 0; f < 6; ++f) {
          bool bnd;
          const int64_t nb = neighbour(i, j, k, f, NeX, NeY, NeZ, bnd);
          for (int fp = 0; fp < Nfp; ++fp) {
            const int a = fp % Nq, b = fp / Nq;
            const int64_t r = (int64_t)f * Nfp + fp + e * NfpTot;
            Fscale[r] = geo[kGeoFscale + f];
            for (int d = 0; d < 3; ++d)
              normal_fn[r + (int64_t)d * NfpTot * Ne] = geo[kGeoNormal + 3 * f + d];
            VMapM[r] = (int)(e * Np + own_face_node(f, a, b, Nq) + 1);
            const int64_t src = nb * Np + nbr_face_node(f, a, b, Nq) + 1;
            if (bnd) {
              VMapP[r] = (int)((int64_t)Np * Ne + ibuf + 1);
              halo_src_map[ibuf++] = (int)src;
            } else {
              VMapP[r] = (int)src;
            }
          }
        }
      }
  return DGK_OK;
}
