/*
 * vfr_host.cpp -- host-side pieces of libvfr.so that need no GPU:
 * the VF-Ray text formats, vertex normalization, limits, the psi table from
 * its closed form, and the seeded synthetic workloads.
 *
 * Built with -ffp-contract=off and without fast-math so that the float
 * results (normalization, limits) are the IEEE results of the reference's
 * operation order.
 */
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vfr.h"
#include "vfr_internal.h"

/* ------------------------------------------------------------------ errors */

static thread_local std::string g_err;

void vfr_set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

extern "C" const char* vfr_last_error(void) { return g_err.c_str(); }

extern "C" void vfr_free(void* p) { free(p); }

/* ------------------------------------------------------------- text input */

namespace {

/* Whitespace-separated token reader over a whole file (the reference reads
 * every format with operator>>, which skips any whitespace). */
struct Tok {
  std::vector<char> buf;
  size_t pos = 0;
  bool ok = false;
  explicit Tok(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n < 0) { fclose(f); return; }
    buf.resize((size_t)n + 1);
    size_t got = n ? fread(buf.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    buf[got] = 0;
    buf.resize(got + 1);
    ok = true;
  }
  const char* next() {
    while (pos < buf.size() - 1 && isspace((unsigned char)buf[pos])) pos++;
    if (pos >= buf.size() - 1) return nullptr;
    const char* s = &buf[pos];
    while (pos < buf.size() - 1 && !isspace((unsigned char)buf[pos])) pos++;
    buf[pos] = 0;
    if (pos < buf.size() - 1) pos++;
    return s;
  }
  bool u32v(uint32_t* out) {
    const char* s = next();
    if (!s) return false;
    char* e;
    errno = 0;
    unsigned long long v = strtoull(s, &e, 10);
    if (*e || errno || v > 0xFFFFFFFFull || s[0] == '-') return false;
    *out = (uint32_t)v;
    return true;
  }
  bool f32v(float* out) {
    const char* s = next();
    if (!s) return false;
    char* e;
    float v = strtof(s, &e);
    if (*e) return false;
    *out = v;
    return true;
  }
};

/* ostream << float with the default precision 6 (vfRayPreComp.cc:638, 863). */
inline void put_float(FILE* f, float v) { fprintf(f, "%g", (double)v); }

}  // namespace

/* readOff, vfRayPreComp.cc:29-80.  Unlike the reference it validates that
 * every tet index is < numVerts. */
extern "C" int vfr_read_off(const char* path, float** xyzs, uint32_t* n_verts, uint32_t** tets,
                            uint32_t* n_tets) {
  if (!path || !xyzs || !n_verts || !tets || !n_tets) {
    vfr_set_error("vfr_read_off: null argument");
    return VFR_EINVAL;
  }
  Tok t(path);
  if (!t.ok) { vfr_set_error("cannot open %s", path); return VFR_EIO; }
  uint32_t nv, nt;
  if (!t.u32v(&nv) || !t.u32v(&nt)) { vfr_set_error("%s: bad header", path); return VFR_EIO; }
  float* v = (float*)malloc(sizeof(float) * 4 * (size_t)nv + 16);
  uint32_t* tt = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)nt + 16);
  if (!v || !tt) { free(v); free(tt); vfr_set_error("out of memory"); return VFR_ENOMEM; }
  for (size_t i = 0; i < 4 * (size_t)nv; ++i)
    if (!t.f32v(&v[i])) { free(v); free(tt); vfr_set_error("%s: bad vertex", path); return VFR_EIO; }
  for (size_t i = 0; i < 4 * (size_t)nt; ++i)
    if (!t.u32v(&tt[i]) || tt[i] >= nv) {
      free(v); free(tt);
      vfr_set_error("%s: bad tetrahedron %zu", path, i / 4);
      return VFR_EIO;
    }
  *xyzs = v; *n_verts = nv; *tets = tt; *n_tets = nt;
  return VFR_OK;
}

extern "C" int vfr_write_off(const char* path, const float* xyzs, uint32_t n_verts,
                             const uint32_t* tets, uint32_t n_tets) {
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("cannot write %s", path); return VFR_EIO; }
  fprintf(f, "%u %u\n", n_verts, n_tets);
  for (uint32_t i = 0; i < n_verts; ++i)
    fprintf(f, "%.9g %.9g %.9g %.9g\n", (double)xyzs[i * 4], (double)xyzs[i * 4 + 1],
            (double)xyzs[i * 4 + 2], (double)xyzs[i * 4 + 3]);
  for (uint32_t i = 0; i < n_tets; ++i)
    fprintf(f, "%u %u %u %u\n", tets[i * 4], tets[i * 4 + 1], tets[i * 4 + 2], tets[i * 4 + 3]);
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* normalizeVertices, vfRayPreComp.cc:84-145: centre the bounding box at the
 * origin, scale by the largest half extent, scalar to [0,1]. The centre and
 * both scale factors are computed in double and rounded to float, as the
 * reference's `/ 2.0` and `1.0 /` literals make them. */
void vfr_normalize_params(const float mn_in[4], const float mx_in[4], float center[3], float* sc, float* ss) {
  float mn[4], mx[4];
  for (int i = 0; i < 4; ++i) { mn[i] = mn_in[i]; mx[i] = mx_in[i]; }
  for (int i = 0; i < 3; ++i) {
    center[i] = (float)((double)(mn[i] + mx[i]) / 2.0);
    mx[i] -= center[i];
    mn[i] -= center[i];
  }
  float m = (mx[1] > mx[2]) ? mx[1] : mx[2];
  m = (mx[0] > m) ? mx[0] : m;
  *sc = (float)(1.0 / (double)m);
  *ss = (float)(1.0 / (double)(mx[3] - mn[3]));
}

extern "C" void vfr_normalize_vertices(float* v, uint32_t nV) {
  if (!v || nV == 0) return;
  float mn[4], mx[4], center[3];
  for (int i = 0; i < 4; ++i) { mn[i] = v[i]; mx[i] = v[i]; }
  for (uint32_t i = 1; i < nV; ++i)
    for (int j = 0; j < 4; ++j) {
      float x = v[(size_t)i * 4 + j];
      if (x < mn[j]) mn[j] = x;
      if (x > mx[j]) mx[j] = x;
    }
  float sc, ss;
  vfr_normalize_params(mn, mx, center, &sc, &ss);
  for (uint32_t i = 0; i < nV; ++i) {
    float* p = v + (size_t)i * 4;
    p[0] = (p[0] - center[0]) * sc;
    p[1] = (p[1] - center[1]) * sc;
    p[2] = (p[2] - center[2]) * sc;
    p[3] = (p[3] - mn[3]) * ss;
  }
}

/* rotateVolume, vfRayPreComp.cc:886-906: row r of the new vertex is
 * sum_c mv[r*4+c] * (x, y, z, 1)[c], accumulated from 0 in c order; x, y, z
 * are replaced and the scalar kept.  mv is read row-major, so a column-major
 * OpenGL matrix M applies as its transpose. */
extern "C" void vfr_rotate_volume(float* v, uint32_t nV, const float mv[16]) {
  if (!v || !mv) return;
  for (uint32_t i = 0; i < nV; ++i) {
    float* p = v + (size_t)i * 4;
    const float o[4] = {p[0], p[1], p[2], 1.0f};
    float r[3];
    for (int a = 0; a < 3; ++a) {
      float acc = 0.0f;
      for (int c = 0; c < 4; ++c) acc = acc + mv[a * 4 + c] * o[c];
      r[a] = acc;
    }
    p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
  }
}

/* The modelview matrix of the viewer's space key (vfRay.cu:618-622):
 * identity, glRotatef(rotx, 1, 0, 0), glRotatef(roty, 0, 1, 0), read back
 * column-major, i.e. M = Rx(rotx) Ry(roty) with mv[c*4+r] = M[r][c].  The
 * entries are computed in double and rounded (a GL implementation's own
 * rounding of the matrix is not reproduced). */
extern "C" void vfr_gl_rotation(float rotx_deg, float roty_deg, float mv[16]) {
  if (!mv) return;
  const double k = 3.14159265358979323846 / 180.0;
  const double cx = cos((double)rotx_deg * k), sx = sin((double)rotx_deg * k);
  const double cy = cos((double)roty_deg * k), sy = sin((double)roty_deg * k);
  const double M[3][3] = {{cy, 0.0, sy}, {sx * sy, cx, -sx * cy}, {-cx * sy, sx, cx * cy}};
  for (int i = 0; i < 16; ++i) mv[i] = 0.0f;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) mv[c * 4 + r] = (float)M[r][c];
  mv[15] = 1.0f;
}

/* buildLmt, vfRayPreComp.cc:808-853: longest tet edge (float sqrt), and the
 * z range of the vertices starting from maxZ = -1, minZ = 1. */
extern "C" void vfr_build_lmt(const float* v, uint32_t nV, const uint32_t* t, uint32_t nT,
                              float out[3]) {
  static const int E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  float maxEdge = 0.0f;
  for (uint32_t i = 0; i < nT; ++i) {
    const uint32_t* q = t + (size_t)i * 4;
    for (int j = 0; j < 6; ++j) {
      const float* a = v + (size_t)q[E[j][0]] * 4;
      const float* b = v + (size_t)q[E[j][1]] * 4;
      float e0 = a[0] - b[0], e1 = a[1] - b[1], e2 = a[2] - b[2];
      float len = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
      if (len > maxEdge) maxEdge = len;
    }
  }
  float maxZ = -1.0f, minZ = 1.0f;
  for (uint32_t i = 0; i < nV; ++i) {
    float z = v[(size_t)i * 4 + 2];
    if (z > maxZ) maxZ = z;
    if (z < minZ) minZ = z;
  }
  out[0] = maxEdge; out[1] = maxZ; out[2] = minZ;
}

/* buildIncid + writeIncid, vfRayPreComp.cc:219-332: per vertex, the tets that
 * contain it in ascending order (a tet once), then the sorted set of the
 * other vertices of those tets; each list prefixed by its size, ids separated
 * by single spaces, a newline after the last id of a non-empty list. */
extern "C" int vfr_write_incid(const char* path, const uint32_t* tets, uint32_t n_tets,
                               uint32_t n_verts) {
  std::vector<uint32_t> cnt((size_t)n_verts + 1, 0);
  for (uint32_t j = 0; j < n_tets; ++j)
    for (int k = 0; k < 4; ++k) {
      bool dup = false;
      for (int q = 0; q < k; ++q) dup |= tets[(size_t)j * 4 + q] == tets[(size_t)j * 4 + k];
      if (!dup) cnt[tets[(size_t)j * 4 + k] + 1]++;
    }
  for (uint32_t i = 0; i < n_verts; ++i) cnt[i + 1] += cnt[i];
  std::vector<uint32_t> lst(cnt[n_verts]), pos(cnt.begin(), cnt.end() - 1);
  for (uint32_t j = 0; j < n_tets; ++j)
    for (int k = 0; k < 4; ++k) {
      bool dup = false;
      for (int q = 0; q < k; ++q) dup |= tets[(size_t)j * 4 + q] == tets[(size_t)j * 4 + k];
      if (!dup) lst[pos[tets[(size_t)j * 4 + k]]++] = j;
    }
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("cannot write %s", path); return VFR_EIO; }
  fprintf(f, "%u\n", n_verts);
  std::vector<uint32_t> nb;
  for (uint32_t i = 0; i < n_verts; ++i) {
    uint32_t n = cnt[i + 1] - cnt[i];
    fprintf(f, "%u ", n);
    nb.clear();
    for (uint32_t q = 0; q < n; ++q) {
      uint32_t j = lst[cnt[i] + q];
      fprintf(f, q + 1 < n ? "%u " : "%u\n", j);
      int k = 0;
      while (k < 4 && tets[(size_t)j * 4 + k] != i) k++;
      for (int m = 1; m <= 3; ++m) nb.push_back(tets[(size_t)j * 4 + ((m + k) % 4)]);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    fprintf(f, "%zu ", nb.size());
    for (size_t q = 0; q < nb.size(); ++q) fprintf(f, q + 1 < nb.size() ? "%u " : "%u\n", nb[q]);
  }
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* readIncid, vfRayPreComp.cc:153-212: the vertex count (must equal the
 * mesh's), then per vertex the number of incident tets and their ids, the
 * number of neighbouring vertices and their ids.  The tet lists come back as
 * CSR arrays (offsets n_verts + 1, ids); the vertex lists are parsed and
 * dropped, as buildCon does not use them. */
extern "C" int vfr_read_incid(const char* path, uint32_t n_verts, uint32_t** offsets, uint32_t** tet_ids) {
  if (!path || !offsets || !tet_ids) { vfr_set_error("vfr_read_incid: null argument"); return VFR_EINVAL; }
  Tok t(path);
  if (!t.ok) { vfr_set_error("cannot open %s", path); return VFR_EIO; }
  uint32_t n;
  if (!t.u32v(&n)) { vfr_set_error("%s: bad header", path); return VFR_EIO; }
  if (n != n_verts) { vfr_set_error("%s: %u vertices, mesh has %u", path, n, n_verts); return VFR_EIO; }
  std::vector<uint32_t> off((size_t)n_verts + 1, 0), ids;
  for (uint32_t i = 0; i < n_verts; ++i) {
    uint32_t nt, nv, x;
    if (!t.u32v(&nt)) { vfr_set_error("%s: bad tet count of vertex %u", path, i); return VFR_EIO; }
    for (uint32_t j = 0; j < nt; ++j) {
      if (!t.u32v(&x)) { vfr_set_error("%s: bad tet id of vertex %u", path, i); return VFR_EIO; }
      ids.push_back(x);
    }
    off[i + 1] = (uint32_t)ids.size();
    if (!t.u32v(&nv)) { vfr_set_error("%s: bad vertex count of vertex %u", path, i); return VFR_EIO; }
    for (uint32_t j = 0; j < nv; ++j)
      if (!t.u32v(&x)) { vfr_set_error("%s: bad vertex id of vertex %u", path, i); return VFR_EIO; }
  }
  uint32_t* o = (uint32_t*)malloc(sizeof(uint32_t) * off.size());
  uint32_t* d = (uint32_t*)malloc(sizeof(uint32_t) * ids.size() + 16);
  if (!o || !d) { free(o); free(d); vfr_set_error("out of memory"); return VFR_ENOMEM; }
  memcpy(o, off.data(), sizeof(uint32_t) * off.size());
  if (!ids.empty()) memcpy(d, ids.data(), sizeof(uint32_t) * ids.size());
  *offsets = o;
  *tet_ids = d;
  return VFR_OK;
}

/* buildCon, vfRayPreComp.cc:424-501, over the incidence lists of a .incid
 * file: for each tet and face f (vertices a, b, c of slots f, f+1, f+2), the
 * first tet other than itself in a's list that the lists of b and c also
 * hold, each of those two scanned in file order and given up at the first id
 * above the candidate; the tet itself when none (an external face, counted).
 * With the lists buildIncid writes (ascending) this is the smallest other tet
 * holding the face; with other lists it is what the reference computes from
 * them. */
extern "C" int vfr_con_from_incid(const uint32_t* tets, uint32_t n_tets, uint32_t n_verts, const uint32_t* offsets,
                                  const uint32_t* tet_ids, uint32_t* con_out, uint32_t* n_ext_out) {
  if (!tets || !offsets || !con_out || !n_ext_out) { vfr_set_error("vfr_con_from_incid: null argument"); return VFR_EINVAL; }
  uint32_t n_ext = 0;
  for (uint32_t i = 0; i < n_tets; ++i)
    for (uint32_t f = 0; f < 4; ++f) {
      uint32_t v[3];
      for (uint32_t k = 0; k < 3; ++k) {
        v[k] = tets[(size_t)i * 4 + ((k + f) & 3u)];
        if (v[k] >= n_verts) { vfr_set_error("tet %u references vertex %u >= %u", i, v[k], n_verts); return VFR_EINVAL; }
      }
      uint32_t& c = con_out[(size_t)i * 4 + f];
      c = i;
      for (uint32_t j = offsets[v[0]]; j < offsets[v[0] + 1]; ++j) {
        const uint32_t cur = tet_ids[j];
        if (cur == i) continue;
        for (uint32_t k = offsets[v[1]]; k < offsets[v[1] + 1]; ++k) {
          if (tet_ids[k] > cur) break;
          if (tet_ids[k] == cur) {
            for (uint32_t l = offsets[v[2]]; l < offsets[v[2] + 1]; ++l) {
              if (tet_ids[l] > cur) break;
              if (tet_ids[l] == cur) { c = cur; break; }
            }
            if (c == cur) break;
          }
        }
        if (c == cur) break;
      }
      if (c == i) ++n_ext;
    }
  *n_ext_out = n_ext;
  return VFR_OK;
}

/* readCon, vfRayPreComp.cc:375-419: header numExtFaces, numTets (must equal
 * the mesh's), then 4 ids per tet. */
extern "C" int vfr_read_con(const char* path, uint32_t n_tets, uint32_t** con, uint32_t* n_ext) {
  if (!path || !con || !n_ext) { vfr_set_error("vfr_read_con: null argument"); return VFR_EINVAL; }
  Tok t(path);
  if (!t.ok) { vfr_set_error("cannot open %s", path); return VFR_EIO; }
  uint32_t ne, nc;
  if (!t.u32v(&ne) || !t.u32v(&nc)) { vfr_set_error("%s: bad header", path); return VFR_EIO; }
  if (nc != n_tets) { vfr_set_error("%s: %u tets, mesh has %u", path, nc, n_tets); return VFR_EIO; }
  uint32_t* c = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)n_tets + 16);
  if (!c) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  for (size_t i = 0; i < 4 * (size_t)n_tets; ++i)
    if (!t.u32v(&c[i]) || c[i] >= n_tets) {
      free(c);
      vfr_set_error("%s: bad entry %zu", path, i);
      return VFR_EIO;
    }
  *con = c; *n_ext = ne;
  return VFR_OK;
}

/* writeCon, vfRayPreComp.cc:507-543. */
extern "C" int vfr_write_con(const char* path, const uint32_t* con, uint32_t n_tets, uint32_t n_ext) {
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("cannot write %s", path); return VFR_EIO; }
  fprintf(f, "%u\n%u\n", n_ext, n_tets);
  for (uint32_t i = 0; i < n_tets; ++i)
    fprintf(f, "%u %u %u %u\n", con[i * 4], con[i * 4 + 1], con[i * 4 + 2], con[i * 4 + 3]);
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* readExtF, vfRayPreComp.cc:668-705: count must equal numExtFaces from .con. */
extern "C" int vfr_read_extf(const char* path, uint32_t n_ext, uint32_t** extf) {
  if (!path || !extf) { vfr_set_error("vfr_read_extf: null argument"); return VFR_EINVAL; }
  Tok t(path);
  if (!t.ok) { vfr_set_error("cannot open %s", path); return VFR_EIO; }
  uint32_t n;
  if (!t.u32v(&n)) { vfr_set_error("%s: bad header", path); return VFR_EIO; }
  if (n != n_ext) { vfr_set_error("%s: %u faces, .con says %u", path, n, n_ext); return VFR_EIO; }
  uint32_t* e = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)n + 16);
  if (!e) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  for (size_t i = 0; i < 2 * (size_t)n; ++i)
    if (!t.u32v(&e[i]) || ((i & 1) && e[i] > 3)) {
      free(e);
      vfr_set_error("%s: bad entry %zu", path, i / 2);
      return VFR_EIO;
    }
  *extf = e;
  return VFR_OK;
}

/* writeExtF, vfRayPreComp.cc:743-774. */
extern "C" int vfr_write_extf(const char* path, const uint32_t* extf, uint32_t n_ext) {
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("cannot write %s", path); return VFR_EIO; }
  fprintf(f, "%u\n", n_ext);
  for (uint32_t i = 0; i < n_ext; ++i) fprintf(f, "%u %u\n", extf[i * 2], extf[i * 2 + 1]);
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* readTF, vfRayPreComp.cc:551-590: exactly 256 entries. */
extern "C" int vfr_read_tf(const char* path, float rgba[1024]) {
  if (!path || !rgba) { vfr_set_error("vfr_read_tf: null argument"); return VFR_EINVAL; }
  Tok t(path);
  if (!t.ok) { vfr_set_error("cannot open %s", path); return VFR_EIO; }
  uint32_t n;
  if (!t.u32v(&n) || n != 256) { vfr_set_error("%s: need 256 entries", path); return VFR_EIO; }
  float tmp[1024];
  for (int i = 0; i < 1024; ++i)
    if (!t.f32v(&tmp[i])) { vfr_set_error("%s: bad entry %d", path, i / 4); return VFR_EIO; }
  memcpy(rgba, tmp, sizeof(tmp));
  return VFR_OK;
}

/* writeTF, vfRayPreComp.cc:628-660. */
extern "C" int vfr_write_tf(const char* path, const float rgba[1024]) {
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("cannot write %s", path); return VFR_EIO; }
  fprintf(f, "256\n");
  for (int i = 0; i < 256; ++i) {
    for (int c = 0; c < 4; ++c) {
      if (c) fputc(' ', f);
      put_float(f, rgba[i * 4 + c]);
    }
    fputc('\n', f);
  }
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* cleanTF, cleanTF.cc:14-78: a legacy transfer-function file (7 header
 * lines, each read into a 200-character buffer, then 256 lines
 * "index r g b a") rewritten as the .tf text: "256", then "r g b a" per
 * entry, printed from doubles with 6 significant digits. */
extern "C" int vfr_clean_tf(const char* in_path, const char* out_path) {
  if (!in_path || !out_path) { vfr_set_error("vfr_clean_tf: null argument"); return VFR_EINVAL; }
  FILE* f = fopen(in_path, "rb");
  if (!f) { vfr_set_error("cannot open %s", in_path); return VFR_EIO; }
  for (int l = 0; l < 7; ++l) {
    /* a header line of 200 characters or more fails the read */
    int c, len = 0;
    while ((c = fgetc(f)) != EOF && c != '\n') ++len;
    if (len >= 200 || (c == EOF && l < 6)) {
      fclose(f);
      vfr_set_error("%s: bad header line %d", in_path, l + 1);
      return VFR_EIO;
    }
  }
  double tf[256][4];
  for (int i = 0; i < 256; ++i) {
    int index;
    if (fscanf(f, "%d %lf %lf %lf %lf", &index, &tf[i][0], &tf[i][1], &tf[i][2], &tf[i][3]) != 5) {
      fclose(f);
      vfr_set_error("%s: bad entry %d", in_path, i);
      return VFR_EIO;
    }
  }
  fclose(f);
  FILE* o = fopen(out_path, "w");
  if (!o) { vfr_set_error("cannot write %s", out_path); return VFR_EIO; }
  fprintf(o, "256\n");
  for (int i = 0; i < 256; ++i) fprintf(o, "%g %g %g %g\n", tf[i][0], tf[i][1], tf[i][2], tf[i][3]);
  return fclose(o) == 0 ? VFR_OK : VFR_EIO;
}

/* buildTF, vfRayPreComp.cc:595-622: four bands of 64 entries (red, yellow,
 * cyan, blue), alpha 0.2. */
extern "C" void vfr_default_tf(float rgba[1024]) {
  static const float band[4][4] = {{1, 0, 0, 0.2f}, {1, 1, 0, 0.2f}, {0, 1, 1, 0.2f}, {0, 0, 1, 0.2f}};
  for (int i = 0; i < 256; ++i)
    for (int c = 0; c < 4; ++c) rgba[i * 4 + c] = band[i / 64][c];
}

/* readLmt / writeLmt, vfRayPreComp.cc:782-880 (6 significant digits). */
extern "C" int vfr_read_lmt(const char* path, float lmt[3]) {
  if (!path || !lmt) { vfr_set_error("vfr_read_lmt: null argument"); return VFR_EINVAL; }
  Tok t(path);
  if (!t.ok) { vfr_set_error("cannot open %s", path); return VFR_EIO; }
  float a, b, c;
  if (!t.f32v(&a) || !t.f32v(&b) || !t.f32v(&c)) { vfr_set_error("%s: malformed", path); return VFR_EIO; }
  lmt[0] = a; lmt[1] = b; lmt[2] = c;
  return VFR_OK;
}

extern "C" int vfr_write_lmt(const char* path, const float lmt[3]) {
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("cannot write %s", path); return VFR_EIO; }
  put_float(f, lmt[0]); fputc(' ', f);
  put_float(f, lmt[1]); fputc(' ', f);
  put_float(f, lmt[2]); fputc('\n', f);
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* ppmWrite, vfRayPreComp.cc:915-934: P6, rows in pix.y order. */
extern "C" int vfr_write_ppm(const char* path, const uint8_t* rgb, uint32_t w, uint32_t h) {
  FILE* f = fopen(path, "w");
  if (!f) { vfr_set_error("Could not save ppm file %s", path); return VFR_EIO; }
  fprintf(f, "P6\n%u %u\n255\n", w, h);
  fwrite(rgb, 1, (size_t)w * h * 3, f);
  return fclose(f) == 0 ? VFR_OK : VFR_EIO;
}

/* ---------------------------------------------------------------- psi table */

/* Psi(g_f, g_b) = int_0^1 exp(-tf s - (tb - tf) s^2 / 2) ds with
 * tau = gamma / (1 - gamma), gamma_i = i / 512 (row = front, column = back),
 * the integral psiGammaTable512.h tabulates.  Composite 8-point
 * Gauss-Legendre in double, one panel per unit of decay rate. */
static const double kGLX[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                               -0.1834346424956498, 0.1834346424956498, 0.5255324099163290,
                               0.7966664774136267, 0.9602898564975363};
static const double kGLW[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                               0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                               0.2223810344533745, 0.1012285362903763};

double vfr_psi_integral(double tf, double tb) {
  double m = tf > tb ? tf : tb;
  int panels = 1 + (int)(m * 0.5);
  if (panels > 4096) panels = 4096;
  double h = 1.0 / panels, sum = 0.0;
  for (int p = 0; p < panels; ++p) {
    double lo = p * h;
    for (int k = 0; k < 8; ++k) {
      double s = lo + 0.5 * h * (kGLX[k] + 1.0);
      sum += kGLW[k] * exp(-tf * s - (tb - tf) * s * s * 0.5);
    }
  }
  return sum * 0.5 * h;
}

extern "C" int vfr_psi_generate(float* out) {
  if (!out) { vfr_set_error("vfr_psi_generate: null argument"); return VFR_EINVAL; }
  double tau[512];
  for (int i = 0; i < 512; ++i) {
    double g = i / 512.0;
    tau[i] = g / (1.0 - g);
  }
  for (int i = 0; i < 512; ++i)
    for (int j = 0; j < 512; ++j) out[i * 512 + j] = (float)vfr_psi_integral(tau[i], tau[j]);
  return VFR_OK;
}

/* --------------------------------------------------------------- synthetic */

namespace {

/* splitmix64 (Steele, Lea, Flood 2014): one portable stream per role. */
struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t role) : s(seed ^ (0xD1B54A32D192ED03ull * (role + 1))) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

enum { ROLE_JITTER = 1, ROLE_FIELD = 2, ROLE_TF = 3, ROLE_VOIDS = 4, ROLE_NOISE = 5 };

}  // namespace

/* Hex grid of n^3 cells on [0,1]^3, vertex (i,j,k) has id (k*(n+1)+j)*(n+1)+i
 * and is jittered by U(-0.15h, 0.15h) per coordinate (h = 1/n), tangentially
 * only on the boundary.  Each cell (id (k*n+j)*n+i) is split into 6
 * Freudenthal tets along its 000->111 diagonal, one per axis order
 * xyz, xzy, yxz, yzx, zxy, zyx; odd orders swap slots 1 and 2 so that every
 * tet has positive volume.  Scalar: sum of 4 (variant 0) or 6 (variant 1)
 * isotropic Gaussians with centres U[0.2,0.8]^3 and sigma U[0.08,0.2],
 * variant 1 adds 0.05*U[0,1] noise; normalized to [0,1].  Variant 1 also maps
 * p -> c + (p-c)|p-c|^0.5 (c = centre) and removes the cells whose unmapped
 * centre lies in one of 12 voids (centre U[0.15,0.85]^3, radius
 * U[0.04,0.1]); unreferenced vertices are dropped. */
extern "C" int vfr_synth_grid(uint32_t n, uint64_t seed, uint32_t variant, float** xyzs_out,
                              uint32_t* n_verts, uint32_t** tets_out, uint32_t* n_tets) {
  if (n < 1 || n > 400 || variant > 1 || !xyzs_out || !n_verts || !tets_out || !n_tets) {
    vfr_set_error("vfr_synth_grid: bad argument");
    return VFR_EINVAL;
  }
  const uint32_t m = n + 1;
  const size_t nv = (size_t)m * m * m;
  const double h = 1.0 / n;
  std::vector<double> P(nv * 3);
  Rng rj(seed, ROLE_JITTER);
  for (uint32_t k = 0; k < m; ++k)
    for (uint32_t j = 0; j < m; ++j)
      for (uint32_t i = 0; i < m; ++i) {
        size_t id = ((size_t)k * m + j) * m + i;
        uint32_t idx[3] = {i, j, k};
        for (int a = 0; a < 3; ++a) {
          double u = rj.uni();
          double jit = (idx[a] == 0 || idx[a] == n) ? 0.0 : (2.0 * u - 1.0) * 0.15 * h;
          P[id * 3 + a] = idx[a] * h + jit;
        }
      }
  if (variant == 1)
    for (size_t id = 0; id < nv; ++id) {
      double d[3], r2 = 0;
      for (int a = 0; a < 3; ++a) { d[a] = P[id * 3 + a] - 0.5; r2 += d[a] * d[a]; }
      double sc = pow(r2, 0.25);
      for (int a = 0; a < 3; ++a) P[id * 3 + a] = 0.5 + d[a] * sc;
    }
  /* scalar field */
  const int ng = variant == 1 ? 6 : 4;
  double gc[6][3], gs[6];
  Rng rf(seed, ROLE_FIELD);
  for (int g = 0; g < ng; ++g) {
    for (int a = 0; a < 3; ++a) gc[g][a] = 0.2 + 0.6 * rf.uni();
    gs[g] = 0.08 + 0.12 * rf.uni();
  }
  std::vector<double> S(nv);
  Rng rn(seed, ROLE_NOISE);
  for (size_t id = 0; id < nv; ++id) {
    double s = 0;
    for (int g = 0; g < ng; ++g) {
      double r2 = 0;
      for (int a = 0; a < 3; ++a) { double d = P[id * 3 + a] - gc[g][a]; r2 += d * d; }
      s += exp(-r2 / (2.0 * gs[g] * gs[g]));
    }
    if (variant == 1) s += 0.05 * rn.uni();
    S[id] = s;
  }
  /* voids */
  double vc[12][3], vr[12];
  if (variant == 1) {
    Rng rv(seed, ROLE_VOIDS);
    for (int q = 0; q < 12; ++q) {
      for (int a = 0; a < 3; ++a) vc[q][a] = 0.15 + 0.7 * rv.uni();
      vr[q] = 0.04 + 0.06 * rv.uni();
    }
  }
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const bool odd[6] = {false, true, true, false, false, true};
  std::vector<uint32_t> T;
  T.reserve((size_t)n * n * n * 24);
  for (uint32_t k = 0; k < n; ++k)
    for (uint32_t j = 0; j < n; ++j)
      for (uint32_t i = 0; i < n; ++i) {
        if (variant == 1) {
          double c[3] = {(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h};
          bool hole = false;
          for (int q = 0; q < 12 && !hole; ++q) {
            double r2 = 0;
            for (int a = 0; a < 3; ++a) { double d = c[a] - vc[q][a]; r2 += d * d; }
            hole = r2 < vr[q] * vr[q];
          }
          if (hole) continue;
        }
        for (int p = 0; p < 6; ++p) {
          uint32_t c[3] = {i, j, k};
          uint32_t v[4];
          v[0] = (c[2] * m + c[1]) * m + c[0];
          c[perm[p][0]]++;
          v[1] = (c[2] * m + c[1]) * m + c[0];
          c[perm[p][1]]++;
          v[2] = (c[2] * m + c[1]) * m + c[0];
          c[perm[p][2]]++;
          v[3] = (c[2] * m + c[1]) * m + c[0];
          if (odd[p]) std::swap(v[1], v[2]);
          for (int q = 0; q < 4; ++q) T.push_back(v[q]);
        }
      }
  /* drop unreferenced vertices (keeps order) */
  std::vector<uint32_t> remap(nv, 0xFFFFFFFFu);
  for (uint32_t id : T) remap[id] = 0;
  uint32_t cntv = 0;
  for (size_t id = 0; id < nv; ++id)
    if (remap[id] == 0) remap[id] = cntv++;
  double smin = 1e300, smax = -1e300;
  for (size_t id = 0; id < nv; ++id)
    if (remap[id] != 0xFFFFFFFFu) { smin = std::min(smin, S[id]); smax = std::max(smax, S[id]); }
  float* X = (float*)malloc(sizeof(float) * 4 * (size_t)cntv + 16);
  uint32_t* TT = (uint32_t*)malloc(sizeof(uint32_t) * T.size() + 16);
  if (!X || !TT) { free(X); free(TT); vfr_set_error("out of memory"); return VFR_ENOMEM; }
  for (size_t id = 0; id < nv; ++id) {
    uint32_t r = remap[id];
    if (r == 0xFFFFFFFFu) continue;
    for (int a = 0; a < 3; ++a) X[(size_t)r * 4 + a] = (float)P[id * 3 + a];
    X[(size_t)r * 4 + 3] = (float)((S[id] - smin) / (smax - smin));
  }
  for (size_t q = 0; q < T.size(); ++q) TT[q] = remap[T[q]];
  /* every tet must keep a positive volume after jitter and mapping */
  for (size_t t = 0; t < T.size() / 4; ++t) {
    const float* a = X + (size_t)TT[t * 4] * 4;
    const float* b = X + (size_t)TT[t * 4 + 1] * 4;
    const float* c = X + (size_t)TT[t * 4 + 2] * 4;
    const float* d = X + (size_t)TT[t * 4 + 3] * 4;
    double e1[3], e2[3], e3[3];
    for (int q = 0; q < 3; ++q) { e1[q] = b[q] - a[q]; e2[q] = c[q] - a[q]; e3[q] = d[q] - a[q]; }
    double vol = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                 e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    if (!(vol > 0.0)) {
      free(X); free(TT);
      vfr_set_error("vfr_synth_grid: tet %zu has volume %g", t, vol);
      return VFR_EMESH;
    }
  }
  *xyzs_out = X; *n_verts = cntv; *tets_out = TT; *n_tets = (uint32_t)(T.size() / 4);
  return VFR_OK;
}

/* 6 control points at x = 0, four sorted U[0,1], 1; rgb U[0,1], alpha
 * U[0,0.4]; linear between them, sampled at i/255, then passed through the
 * .tf text form (6 significant digits) so a file round trip is exact. */
extern "C" void vfr_synth_tf(uint64_t seed, float rgba[1024]) {
  Rng r(seed, ROLE_TF);
  double x[6], c[6][4];
  x[0] = 0.0; x[5] = 1.0;
  for (int i = 1; i < 5; ++i) x[i] = r.uni();
  std::sort(x + 1, x + 5);
  for (int i = 0; i < 6; ++i) {
    for (int q = 0; q < 3; ++q) c[i][q] = r.uni();
    c[i][3] = 0.4 * r.uni();
  }
  for (int e = 0; e < 256; ++e) {
    double u = e / 255.0;
    int s = 0;
    while (s < 4 && u > x[s + 1]) s++;
    double w = (x[s + 1] > x[s]) ? (u - x[s]) / (x[s + 1] - x[s]) : 0.0;
    for (int q = 0; q < 4; ++q) {
      double val = c[s][q] + w * (c[s + 1][q] - c[s][q]);
      char buf[64];
      snprintf(buf, sizeof(buf), "%g", (double)(float)val);
      rgba[e * 4 + q] = strtof(buf, nullptr);
    }
  }
}

/* Rotation R with R d = (0,0,-1), applied about (0.5,0.5,0.5). */
extern "C" void vfr_synth_rotate(float* v, uint32_t nV, const float d_in[3]) {
  double d[3] = {d_in[0], d_in[1], d_in[2]};
  double ln = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int a = 0; a < 3; ++a) d[a] /= ln;
  double r2[3] = {-d[0], -d[1], -d[2]};
  double up[3] = {0, 0, 0};
  int ax = 0;
  for (int a = 1; a < 3; ++a)
    if (fabs(d[a]) < fabs(d[ax])) ax = a;
  up[ax] = 1.0;
  double r0[3] = {up[1] * r2[2] - up[2] * r2[1], up[2] * r2[0] - up[0] * r2[2],
                  up[0] * r2[1] - up[1] * r2[0]};
  double l0 = sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
  for (int a = 0; a < 3; ++a) r0[a] /= l0;
  double r1[3] = {r2[1] * r0[2] - r2[2] * r0[1], r2[2] * r0[0] - r2[0] * r0[2],
                  r2[0] * r0[1] - r2[1] * r0[0]};
  for (uint32_t i = 0; i < nV; ++i) {
    float* p = v + (size_t)i * 4;
    double q[3] = {p[0] - 0.5, p[1] - 0.5, p[2] - 0.5};
    p[0] = (float)(r0[0] * q[0] + r0[1] * q[1] + r0[2] * q[2]);
    p[1] = (float)(r1[0] * q[0] + r1[1] * q[1] + r1[2] * q[2]);
    p[2] = (float)(r2[0] * q[0] + r2[1] * q[1] + r2[2] * q[2]);
  }
}
