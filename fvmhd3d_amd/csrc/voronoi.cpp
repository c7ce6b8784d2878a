// Periodic 3-D Voronoi tessellation by half-space clipping (host C++, OpenMP).
//
// This is the mesh side of the boundary: the reference builds its Voronoi mesh
// with CGAL (localMesh.cpp:333-614), which is not part of the accelerated path
// and is not available here. This builder produces the same face/cell records
// with the same geometric conventions, so the hot-path kernels and the oracle
// receive input exactly like the reference's localMesh_build() hands them over:
//
//  * face polygon fan centre c = mean of the polygon vertices; vector area
//    n = 1/2 sum (v_k - c) x (v_{k+1} - c); centroid = area-weighted average of
//    the fan-triangle centroids (localMesh.cpp:520-543, 556);
//  * faces with sqrt(sum |(v_k-c)x(v_{k+1}-c)|) < 1e-10 |centroid - x_i| are
//    dropped (localMesh.cpp:545-549);
//  * n is oriented from s1 to s2 (localMesh.cpp:558-559), s1 is the cell that
//    creates the face, i.e. the lower cell index unless the lower cell did not
//    see the face (localMesh.cpp:497-591);
//  * every cell's face list is in face-creation order, which equals ascending
//    face index (cell_list[i].ngb order of localMesh.cpp:587-590);
//  * cell volume and centroid come from the tetrahedron fan site -> face
//    centroid -> polygon edge (localMesh.cpp:593-611).
//
// Periodicity: a face between i and a periodic image of j is stored with
// s2 = j (the owned index) and with its centroid in the frame of s1, which is
// what the reference's per-face code (dri = centroid - pos[s1]) expects; the
// reconstruction applies the minimum-image wrap itself (computeReconstruction.cpp:97-103).
//
// Each cell's volume uses its own polygons and its own face centroids (the
// reference uses the shared face centroid of the creating cell for the other
// cell, which differs only by rounding).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// Candidate neighbour: owned index j plus the periodic image shift.
struct Cand {
  double d2;
  V3 q;  // position of the image relative to the site
  int j;
  int8_t sx, sy, sz;
};

// Convex polyhedron in coordinates relative to the site.
struct Poly {
  std::vector<V3> v;
  std::vector<int> fstart, fsize, fv;  // face f: fv[fstart[f] .. fstart[f]+fsize[f])
  std::vector<int> label;              // neighbour candidate id, -1 for the initial box
  // scratch
  std::vector<double> d;
  std::vector<int> remap;
  std::vector<V3> nv;
  std::vector<int> nfstart, nfsize, nfv, nlabel;
  std::vector<std::pair<int64_t, int>> xkeys;
  std::vector<std::pair<int, int>> seg;  // (entry, exit) on the cut plane

  void init_box(double h) {
    v.clear();
    for (int k = 0; k < 8; k++)
      v.push_back({(k & 1) ? h : -h, (k & 2) ? h : -h, (k & 4) ? h : -h});
    // outward counter-clockwise faces
    static const int F[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4},
                                {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
    fstart.clear(); fsize.clear(); fv.clear(); label.clear();
    for (int f = 0; f < 6; f++) {
      fstart.push_back((int)fv.size());
      fsize.push_back(4);
      for (int k = 0; k < 4; k++) fv.push_back(F[f][k]);
      label.push_back(-1);
    }
  }

  double rmax2() const {
    double r = 0;
    for (const V3 &p : v) r = std::max(r, dot(p, p));
    return r;
  }

  int xvertex(int a, int b) {
    const int lo = std::min(a, b), hi = std::max(a, b);
    const int64_t key = ((int64_t)lo << 32) | (uint32_t)hi;
    for (auto &kv : xkeys)
      if (kv.first == key) return kv.second;
    const double t = d[lo] / (d[lo] - d[hi]);
    const V3 p = v[lo] + (v[hi] - v[lo]) * t;
    nv.push_back(p);
    xkeys.push_back({key, (int)nv.size() - 1});
    return (int)nv.size() - 1;
  }

  // Clip by the bisector half-space of the image at q. Returns false on a
  // topology failure (the cut did not close into a single cycle).
  bool clip(V3 q, int lab) {
    const double h = 0.5 * dot(q, q);
    const int n = (int)v.size();
    d.resize(n);
    bool any_out = false;
    for (int k = 0; k < n; k++) {
      d[k] = dot(v[k], q) - h;
      if (d[k] > 0) any_out = true;
    }
    if (!any_out) return true;
    remap.assign(n, -1);
    nv.clear();
    for (int k = 0; k < n; k++)
      if (d[k] <= 0) {
        remap[k] = (int)nv.size();
        nv.push_back(v[k]);
      }
    xkeys.clear();
    seg.clear();
    nfstart.clear(); nfsize.clear(); nfv.clear(); nlabel.clear();
    const int nf = (int)fsize.size();
    for (int f = 0; f < nf; f++) {
      const int s = fstart[f], m = fsize[f];
      const int start = (int)nfv.size();
      int pending_exit = -1, first_entry = -1;
      for (int e = 0; e < m; e++) {
        const int a = fv[s + e], b = fv[s + (e + 1) % m];
        const bool ain = d[a] <= 0, bin = d[b] <= 0;
        if (ain) nfv.push_back(remap[a]);
        if (ain != bin) {
          const int x = xvertex(a, b);
          nfv.push_back(x);
          if (ain) {
            pending_exit = x;  // leaving the kept region
          } else {
            if (pending_exit >= 0) {
              seg.push_back({x, pending_exit});
              pending_exit = -1;
            } else {
              first_entry = x;
            }
          }
        }
      }
      if (pending_exit >= 0) {
        if (first_entry < 0) return false;
        seg.push_back({first_entry, pending_exit});
      }
      const int cnt = (int)nfv.size() - start;
      if (cnt >= 3) {
        nfstart.push_back(start);
        nfsize.push_back(cnt);
        nlabel.push_back(label[f]);
      } else {
        nfv.resize(start);
      }
    }
    // chain the cut face: entry -> exit of each clipped face
    if (seg.size() >= 3) {
      const int start = (int)nfv.size();
      int cur = seg[0].first;
      const int first = cur;
      size_t steps = 0;
      do {
        nfv.push_back(cur);
        int nxt = -1;
        for (auto &sg : seg)
          if (sg.first == cur) {
            nxt = sg.second;
            break;
          }
        if (nxt < 0) return false;
        cur = nxt;
        if (++steps > seg.size()) return false;
      } while (cur != first);
      if (steps != seg.size()) return false;
      nfstart.push_back(start);
      nfsize.push_back((int)nfv.size() - start);
      nlabel.push_back(lab);
    }
    v.swap(nv);
    fstart.swap(nfstart);
    fsize.swap(nfsize);
    fv.swap(nfv);
    label.swap(nlabel);
    return true;
  }
};

struct Grid {
  int nb[3];
  double L[3], hb[3];
  std::vector<int> off, idx;
  void build(int n, const double *pos, const double *box) {
    const double vol = box[0] * box[1] * box[2];
    const double h = std::cbrt(vol / std::max(1, n) * 2.0);
    for (int d = 0; d < 3; d++) {
      L[d] = box[d];
      nb[d] = std::max(1, (int)std::floor(box[d] / h));
      hb[d] = box[d] / nb[d];
    }
    const long nbin = (long)nb[0] * nb[1] * nb[2];
    off.assign(nbin + 1, 0);
    std::vector<int> bin(n);
    for (int i = 0; i < n; i++) {
      int b[3];
      for (int d = 0; d < 3; d++) {
        double x = pos[3 * i + d];
        int k = (int)std::floor(x / hb[d]);
        k = ((k % nb[d]) + nb[d]) % nb[d];
        b[d] = k;
      }
      bin[i] = (b[2] * nb[1] + b[1]) * nb[0] + b[0];
      off[bin[i] + 1]++;
    }
    for (long k = 0; k < nbin; k++) off[k + 1] += off[k];
    idx.resize(n);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int i = 0; i < n; i++) idx[fill[bin[i]]++] = i;
  }
};

struct FaceEnd {
  int j;
  int8_t sx, sy, sz;
  int8_t culled;
};

struct Builder {
  int n;
  const double *pos;
  double L[3];
  Grid grid;
  std::string err;

  // candidates in the bin shell [r0, r1] around site i (Chebyshev bin rings)
  void gather(int i, int r0, int r1, std::vector<Cand> &out) const {
    const V3 p = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int b[3];
    const double pp[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; d++) {
      int k = (int)std::floor(pp[d] / grid.hb[d]);
      b[d] = ((k % grid.nb[d]) + grid.nb[d]) % grid.nb[d];
    }
    for (int dz = -r1; dz <= r1; dz++)
      for (int dy = -r1; dy <= r1; dy++)
        for (int dx = -r1; dx <= r1; dx++) {
          const int cheb = std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz)));
          if (cheb < r0) continue;
          int c[3] = {b[0] + dx, b[1] + dy, b[2] + dz};
          int s[3];
          for (int d = 0; d < 3; d++) {
            s[d] = (int)std::floor((double)c[d] / grid.nb[d]);
            c[d] -= s[d] * grid.nb[d];
          }
          const int bin = (c[2] * grid.nb[1] + c[1]) * grid.nb[0] + c[0];
          for (int k = grid.off[bin]; k < grid.off[bin + 1]; k++) {
            const int j = grid.idx[k];
            if (j == i && s[0] == 0 && s[1] == 0 && s[2] == 0) continue;
            const V3 q = {pos[3 * j] + s[0] * L[0] - p.x, pos[3 * j + 1] + s[1] * L[1] - p.y,
                          pos[3 * j + 2] + s[2] * L[2] - p.z};
            out.push_back({dot(q, q), q, j, (int8_t)s[0], (int8_t)s[1], (int8_t)s[2]});
          }
        }
  }

  // Build the Voronoi cell of site i. Returns false on failure.
  bool cell(int i, Poly &P, std::vector<Cand> &cand) const {
    cand.clear();
    double hmin = std::min(grid.hb[0], std::min(grid.hb[1], grid.hb[2]));
    double hmax = std::max(L[0], std::max(L[1], L[2]));
    P.init_box(2.0 * hmax);
    int r = 2;
    gather(i, 0, r, cand);
    size_t done = 0;
    for (;;) {
      std::sort(cand.begin() + done, cand.end(),
                [](const Cand &a, const Cand &b) { return a.d2 < b.d2; });
      // every image closer than `complete` is among the candidates
      const double complete2 = (r * hmin) * (r * hmin);
      size_t k = done;
      for (; k < cand.size(); k++) {
        if (cand[k].d2 >= complete2) break;
        if (cand[k].d2 > 4.0 * P.rmax2()) break;
        if (!P.clip(cand[k].q, (int)k)) return false;
      }
      done = k;
      // security radius: no image beyond 2*rmax can cut the cell
      if (4.0 * P.rmax2() <= complete2) return true;
      if (r + 1 > 64) return false;
      r = r + 1;
      gather(i, r, r, cand);
    }
  }
};

}  // namespace

extern "C" {

typedef struct {
  int s1, s2;
  double centroid[3];
  double n[3];
} fvm_face_rec;

struct fvm_mesh {
  int n;
  std::vector<fvm_face_rec> faces;
  std::vector<int> cell_off, cell_face;
  std::vector<double> volume, centroid;
  std::vector<int> nfaces_poly;  // polygon faces per cell (before dedup)
};

static thread_local std::string g_err;

const char *fvm_mesh_error(void) { return g_err.c_str(); }

// Reference face geometry (localMesh.cpp:520-559) for polygon vtx (absolute
// coordinates in the frame of site ipos). Returns the area1 sum and fills n,
// centroid; culled when the face is dropped by the 1e-10 rule.
static void face_geometry(const std::vector<V3> &vtx, V3 ipos, V3 &n, V3 &cen, bool &culled) {
  const int nvtx = (int)vtx.size();
  V3 c = {0, 0, 0};
  for (int j = 0; j < nvtx; j++) c = c + vtx[j];
  c = c * (1.0 / (double)nvtx);
  n = {0, 0, 0};
  cen = {0, 0, 0};
  double area1 = 0.0;
  const double third = 1.0 / 3.0;
  V3 v1 = vtx[nvtx - 1] - c;
  for (int j = 0; j < nvtx; j++) {
    const V3 v2 = vtx[j] - c;
    const V3 norm3 = cross(v1, v2);
    const double area3 = norm(norm3);
    const V3 c3 = c + (v1 + v2) * third;
    n = n + norm3;
    cen = cen + c3 * area3;
    area1 += area3;
    v1 = v2;
  }
  const double L1 = std::sqrt(area1);
  const double L2 = norm(cen - ipos);
  const double area = (L1 < 1.0e-10 * L2) ? 0.0 : area1;
  const double nabs = norm(n);
  culled = true;
  if (area > 0.0 && nabs > 0.0) {
    const double iarea = 1.0 / area;
    cen = cen * iarea;
    n = n * 0.5;
    if (dot(cen - ipos, n) < 0.0) n = n * -1.0;
    culled = false;
  }
}

int fvm_mesh_build_periodic(int n, const double *pos, const double *box, int nthreads,
                            fvm_mesh **out) {
  *out = nullptr;
  if (n < 8 || !pos || !box) {
    g_err = "fvm_mesh_build_periodic: need n >= 8 sites and a box";
    return -1;
  }
  for (int i = 0; i < 3 * n; i++) {
    const double x = pos[i];
    if (!(x >= 0.0 && x < box[i % 3])) {
      g_err = "fvm_mesh_build_periodic: site outside [0, L)";
      return -1;
    }
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  Builder B;
  B.n = n;
  B.pos = pos;
  for (int d = 0; d < 3; d++) B.L[d] = box[d];
  B.grid.build(n, pos, box);

  // pass 1: neighbour lists with the cull decision of each face-end
  std::vector<int> nend(n, 0);
  std::vector<std::vector<FaceEnd>> ends(n);
  int fail = -1;
#pragma omp parallel
  {
    Poly P;
    std::vector<Cand> cand;
    std::vector<V3> vtx;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
      if (!B.cell(i, P, cand)) {
#pragma omp critical
        fail = i;
        continue;
      }
      const V3 ip = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      std::vector<FaceEnd> &E = ends[i];
      const int nf = (int)P.fsize.size();
      E.reserve(nf);
      for (int f = 0; f < nf; f++) {
        const int lab = P.label[f];
        if (lab < 0) {
#pragma omp critical
          fail = i;
          break;
        }
        vtx.clear();
        for (int k = 0; k < P.fsize[f]; k++) vtx.push_back(ip + P.v[P.fv[P.fstart[f] + k]]);
        V3 nn, cc;
        bool culled;
        face_geometry(vtx, ip, nn, cc, culled);
        const Cand &c = cand[lab];
        E.push_back({c.j, c.sx, c.sy, c.sz, (int8_t)culled});
      }
      nend[i] = (int)E.size();
    }
  }
  if (fail >= 0) {
    g_err = "fvm_mesh_build_periodic: clipping failed at site " + std::to_string(fail);
    return -2;
  }
  // duplicate neighbour check (periodic box too small for the site count)
  for (int i = 0; i < n; i++) {
    std::vector<int> js;
    for (auto &e : ends[i]) js.push_back(e.j);
    std::sort(js.begin(), js.end());
    for (size_t k = 1; k < js.size(); k++)
      if (js[k] == js[k - 1]) {
        g_err = "fvm_mesh_build_periodic: a site neighbours two images of one site; "
                "the box holds too few sites";
        return -3;
      }
  }
  // which face-ends create a face: j > i, or j < i and j has no kept face with i
  auto has_kept = [&](int j, int i) {
    for (auto &e : ends[j])
      if (e.j == i) return e.culled == 0;
    return false;
  };
  std::vector<int> ncreate(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    int c = 0;
    for (auto &e : ends[i]) {
      if (e.culled) continue;
      if (e.j > i || !has_kept(e.j, i)) c++;
    }
    ncreate[i] = c;
  }
  std::vector<long long> fbase(n + 1, 0);
  for (int i = 0; i < n; i++) fbase[i + 1] = fbase[i] + ncreate[i];
  const long long nface = fbase[n];
  if (nface >= (1LL << 31) - 1) {
    g_err = "fvm_mesh_build_periodic: too many faces for int32 indices";
    return -4;
  }
  fvm_mesh *M = new fvm_mesh;
  M->n = n;
  M->faces.resize(nface);
  M->volume.assign(n, 0.0);
  M->centroid.assign(3 * (size_t)n, 0.0);
  M->nfaces_poly.assign(n, 0);

  // pass 2: face geometry of created faces, cell volumes and centroids
#pragma omp parallel
  {
    Poly P;
    std::vector<Cand> cand;
    std::vector<V3> vtx;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
      if (!B.cell(i, P, cand)) continue;  // cannot happen: pass 1 succeeded
      const V3 ip = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      long long fid = fbase[i];
      double cell_volume = 0.0;
      V3 cell_centroid = {0, 0, 0};
      const int nf = (int)P.fsize.size();
      M->nfaces_poly[i] = nf;
      for (int f = 0; f < nf; f++) {
        const Cand &c = cand[P.label[f]];
        vtx.clear();
        for (int k = 0; k < P.fsize[f]; k++) vtx.push_back(ip + P.v[P.fv[P.fstart[f] + k]]);
        V3 nn, cc;
        bool culled;
        face_geometry(vtx, ip, nn, cc, culled);
        if (culled) continue;
        if (c.j > i || !has_kept(c.j, i)) {
          fvm_face_rec &F = M->faces[fid++];
          F.s1 = i;
          F.s2 = c.j;
          F.centroid[0] = cc.x; F.centroid[1] = cc.y; F.centroid[2] = cc.z;
          F.n[0] = nn.x; F.n[1] = nn.y; F.n[2] = nn.z;
        }
        // tetrahedron fan (localMesh.cpp:593-605)
        const V3 cv = ip - cc;
        V3 v1 = vtx.back() - cc;
        const double fourth = 1.0 / 4.0;
        for (size_t k = 0; k < vtx.size(); k++) {
          const V3 v2 = vtx[k] - cc;
          const V3 c4 = cc + (v1 + v2 + cv) * fourth;
          const double vol4 = std::fabs(dot(cross(v1, v2), cv));
          cell_volume += vol4;
          cell_centroid = cell_centroid + c4 * vol4;
          v1 = v2;
        }
      }
      cell_centroid = cell_centroid * (1.0 / cell_volume);
      cell_volume *= 1.0 / 6.0;
      M->volume[i] = cell_volume;
      M->centroid[3 * (size_t)i + 0] = cell_centroid.x;
      M->centroid[3 * (size_t)i + 1] = cell_centroid.y;
      M->centroid[3 * (size_t)i + 2] = cell_centroid.z;
    }
  }
  // cell face lists in ascending face index
  M->cell_off.assign(n + 1, 0);
  for (long long f = 0; f < nface; f++) {
    M->cell_off[M->faces[f].s1 + 1]++;
    M->cell_off[M->faces[f].s2 + 1]++;
  }
  for (int i = 0; i < n; i++) M->cell_off[i + 1] += M->cell_off[i];
  M->cell_face.resize(M->cell_off[n]);
  std::vector<int> fill(M->cell_off.begin(), M->cell_off.end() - 1);
  for (long long f = 0; f < nface; f++) {
    M->cell_face[fill[M->faces[f].s1]++] = (int)f;
    M->cell_face[fill[M->faces[f].s2]++] = (int)f;
  }
  *out = M;
  return 0;
}

long long fvm_mesh_nface(const fvm_mesh *M) { return M ? (long long)M->faces.size() : -1; }
long long fvm_mesh_nref(const fvm_mesh *M) { return M ? (long long)M->cell_face.size() : -1; }

int fvm_mesh_copy(const fvm_mesh *M, fvm_face_rec *faces, int *cell_off, int *cell_face,
                  double *volume, double *centroid) {
  if (!M) return -1;
  if (faces) std::memcpy(faces, M->faces.data(), M->faces.size() * sizeof(fvm_face_rec));
  if (cell_off) std::memcpy(cell_off, M->cell_off.data(), M->cell_off.size() * sizeof(int));
  if (cell_face) std::memcpy(cell_face, M->cell_face.data(), M->cell_face.size() * sizeof(int));
  if (volume) std::memcpy(volume, M->volume.data(), M->volume.size() * sizeof(double));
  if (centroid) std::memcpy(centroid, M->centroid.data(), M->centroid.size() * sizeof(double));
  return 0;
}

void fvm_mesh_free(fvm_mesh *M) { delete M; }

}  // extern "C"

extern "C" {
// Rand48 stream of the reference (fvmhd3d.h:27-49): srand(seed) then n
// successive drand() draws.
int fvm_rand48_drand(long long seed, long long n, double *out) {
  if (n < 0 || (n > 0 && !out)) return -1;
  unsigned long long stat = ((unsigned long long)seed << 16) + 0x330EULL;
  for (long long k = 0; k < n; k++) {
    stat = stat * 0x5DEECE66DULL + 0xBULL;
    out[k] = (double)(stat & 0xFFFFFFFFFFFFULL) * (1.0 / 281474976710656.0);
  }
  return 0;
}
}
