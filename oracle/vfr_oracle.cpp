/*
 * vfr_oracle.cpp -- CPU oracle for the VF-Ray hot path.  TEST INFRASTRUCTURE.
 *
 * Only the test suite, the smoke check and the benchmark's CPU-baseline leg
 * load this library, and only to check or time the renderer against it: the
 * product path (libvfr.so) never links, loads or calls it.
 *
 * What it restates (file:line in the VF-Ray sources):
 *   orc_normalize        normalizeVertices        vfRayPreComp.cc:84-145
 *   orc_connectivity     buildIncid + buildCon    vfRayPreComp.cc:219-270, 424-501
 *   orc_extf             buildExtF                vfRayPreComp.cc:710-737
 *   orc_lmt              buildLmt                 vfRayPreComp.cc:808-853
 *   orc_render_ref       main.cc executeKernel1..3 (main.cc:74-224) over
 *                        findVisibleFaces         vfRayKernelCPU.cc:51-132
 *                        projectVisibleFaces      vfRayKernelCPU.cc:135-169
 *                        vfRay                    vfRayKernelCPU.cc:172-566
 *   orc_render_ftb       the renderer's own front-to-back mode (perspective,
 *                        early termination, re-entry): no reference
 *                        counterpart; entry faces by the reference's inside
 *                        test (vfRayKernelCPU.cc:256-268), exit faces by 2-D
 *                        edge functions oriented by vertex id (watertight),
 *                        the PPI segment formula (vfRayKernelCPU.cc:448-482)
 *                        with front = entry point.
 *   FP64 re-integration  (both render functions) the same tet sequence
 *                        integrated in double with the exact psi integral.
 *
 * Arithmetic: strict IEEE binary32 in the reference's operation order, with
 * the reference's double-promoted sub-expressions made explicit (the
 * reference declares real = float, RDELTA as the double 1e-12,
 * vfRayKernelCPU.h:28-33).  Build with -ffp-contract=off and without
 * -ffast-math.
 *
 * Pinning (what checks each function; the tests hold the checks):
 *   normalize/connectivity/extf/lmt : byte-equal against the compiled
 *     reference (oracle/_ref) on synthetic meshes; connectivity also against
 *     an O(T^2) brute force on tiny meshes.
 *   orc_render_ref : per-pixel last-writer face ids and uchar image against
 *     the compiled reference built with strict IEEE flags; closed-form
 *     single-tet and constant-TF cases.
 *   orc_render_ftb : entry faces against brute-force ray/triangle tests in
 *     double; colours against the FP64 re-integration and closed forms.
 *   FP64 re-integration : closed forms (constant TF, single tet), psi
 *     integral against scipy quadrature.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

typedef uint32_t u32;
typedef uint64_t u64;

static const double RDELTA = 1e-12; /* vfRayKernelCPU.h:28 */

/* A walk longer than this stops as an anomaly (the reference has no bound; a
 * ray cycling around a degenerate edge would never end). Same bound as the
 * renderer. */
static const u32 MAX_STEPS = 8192u;

struct F4 { float x, y, z, w; };

static inline F4 ld4(const float* v, u32 i) {
  F4 r; r.x = v[i * 4 + 0]; r.y = v[i * 4 + 1]; r.z = v[i * 4 + 2]; r.w = v[i * 4 + 3];
  return r;
}

/* ------------------------------------------------------------ preprocessing */

/* normalizeVertices, vfRayPreComp.cc:84-145. */
/* Threads of the OpenMP loops (the render's pixel loop); returns the
 * previous maximum.  For timing the oracle on one core. */
extern "C" int orc_set_threads(int n) {
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
}

extern "C" void orc_normalize(float* v, u32 nV) {
  float mn[4], mx[4], center[3];
  for (int i = 0; i < 4; ++i) { mn[i] = v[i]; mx[i] = v[i]; }
  for (u32 i = 1; i < nV; ++i)
    for (int j = 0; j < 4; ++j) {
      float value = v[i * 4 + j];
      if (value < mn[j]) mn[j] = value;
      if (value > mx[j]) mx[j] = value;
    }
  for (int i = 0; i < 3; ++i) {
    center[i] = (float)((double)(mn[i] + mx[i]) / 2.0);
    mx[i] -= center[i];
    mn[i] -= center[i];
  }
  float maxCoord = (mx[1] > mx[2]) ? mx[1] : mx[2];
  maxCoord = (mx[0] > maxCoord) ? mx[0] : maxCoord;
  float scaleCoord = (float)(1.0 / (double)maxCoord);
  float scaleScalar = (float)(1.0 / (double)(mx[3] - mn[3]));
  for (u32 i = 0; i < nV; ++i) {
    for (int j = 0; j < 3; ++j) v[i * 4 + j] = (v[i * 4 + j] - center[j]) * scaleCoord;
    v[i * 4 + 3] = (v[i * 4 + 3] - mn[3]) * scaleScalar;
  }
}

/* rotateVolume, vfRayPreComp.cc:886-906: vRot[r] += modelView[r*4+c] *
 * vOrig[c] over r, c = 0..3 from vRot = 0, vOrig = (x, y, z, 1); x, y, z
 * take vRot[0..2]. */
extern "C" void orc_rotate(float* v, u32 nV, const float* mv) {
  for (u32 i = 0; i < nV; ++i) {
    float vOrig[4] = {v[i * 4 + 0], v[i * 4 + 1], v[i * 4 + 2], 1.0f};
    float vRot[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (u32 r = 0; r < 4; ++r)
      for (u32 c = 0; c < 4; ++c) vRot[r] += mv[r * 4 + c] * vOrig[c];
    for (u32 j = 0; j < 3; ++j) v[i * 4 + j] = vRot[j];
  }
}

/* buildIncid's per-vertex ascending tet lists (vfRayPreComp.cc:231-266, a tet
 * is listed once per vertex it contains) followed by buildCon's search
 * (vfRayPreComp.cc:438-497): the neighbour across face f of tet i is the
 * smallest tet id != i that appears in the incidence lists of all three face
 * vertices MOD4(k,f), k=0..2; none -> i itself (external). Returns numExtFaces
 * or -1 when a vertex has no incident tet (buildIncid's anomaly, :260-264). */
extern "C" long orc_connectivity(const u32* tets, u32 nT, u32 nV, u32* con) {
  std::vector<u32> cnt(nV + 1, 0);
  for (u32 j = 0; j < nT; ++j) {
    /* one entry per (vertex, tet): the reference breaks at the first slot k
     * that matches vertex i, so a repeated vertex id lists the tet once */
    for (int k = 0; k < 4; ++k) {
      u32 vi = tets[j * 4 + k];
      bool dup = false;
      for (int q = 0; q < k; ++q) dup |= (tets[j * 4 + q] == vi);
      if (!dup) cnt[vi + 1]++;
    }
  }
  for (u32 i = 0; i < nV; ++i) {
    if (cnt[i + 1] == 0) return -1;
    cnt[i + 1] += cnt[i];
  }
  std::vector<u32> lst(cnt[nV]);
  std::vector<u32> pos(cnt.begin(), cnt.end() - 1);
  for (u32 j = 0; j < nT; ++j)
    for (int k = 0; k < 4; ++k) {
      u32 vi = tets[j * 4 + k];
      bool dup = false;
      for (int q = 0; q < k; ++q) dup |= (tets[j * 4 + q] == vi);
      if (!dup) lst[pos[vi]++] = j;
    }
  long numExt = 0;
  for (u32 i = 0; i < nT; ++i) {
    for (u32 f = 0; f < 4; ++f) {
      u32 cv[3];
      for (u32 k = 0; k < 3; ++k) cv[k] = tets[i * 4 + ((k + f) % 4)];
      con[i * 4 + f] = i;
      const u32* l0 = &lst[cnt[cv[0]]]; u32 n0 = cnt[cv[0] + 1] - cnt[cv[0]];
      const u32* l1 = &lst[cnt[cv[1]]]; u32 n1 = cnt[cv[1] + 1] - cnt[cv[1]];
      const u32* l2 = &lst[cnt[cv[2]]]; u32 n2 = cnt[cv[2] + 1] - cnt[cv[2]];
      for (u32 j = 0; j < n0; ++j) {
        u32 cur = l0[j];
        if (i == cur) continue;
        bool found = false;
        for (u32 k = 0; k < n1; ++k) {
          if (l1[k] > cur) break;
          if (l1[k] == cur) {
            for (u32 l = 0; l < n2; ++l) {
              if (l2[l] > cur) break;
              if (l2[l] == cur) { found = true; break; }
            }
            if (found) break;
          }
        }
        if (found) { con[i * 4 + f] = cur; break; }
      }
      if (con[i * 4 + f] == i) numExt++;
    }
  }
  return numExt;
}

/* buildExtF, vfRayPreComp.cc:710-737. */
extern "C" u32 orc_extf(const u32* con, u32 nT, u32* extf) {
  u32 n = 0;
  for (u32 i = 0; i < nT; ++i)
    for (u32 f = 0; f < 4; ++f)
      if (con[i * 4 + f] == i) {
        if (extf) { extf[n * 2 + 0] = i; extf[n * 2 + 1] = f; }
        n++;
      }
  return n;
}

/* buildLmt, vfRayPreComp.cc:808-853. */
extern "C" void orc_lmt(const float* v, u32 nV, const u32* t, u32 nT, float out[3]) {
  static const int E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  float maxEdge = 0.0f;
  for (u32 i = 0; i < nT; ++i)
    for (int j = 0; j < 6; ++j) {
      float e[3];
      for (int k = 0; k < 3; ++k)
        e[k] = v[t[i * 4 + E[j][0]] * 4 + k] - v[t[i * 4 + E[j][1]] * 4 + k];
      float len = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      if (len > maxEdge) maxEdge = len;
    }
  float maxZ = -1.0f, minZ = 1.0f;
  for (u32 i = 0; i < nV; ++i) {
    if (v[i * 4 + 2] > maxZ) maxZ = v[i * 4 + 2];
    if (v[i * 4 + 2] < minZ) minZ = v[i * 4 + 2];
  }
  out[0] = maxEdge; out[1] = maxZ; out[2] = minZ;
}

/* ----------------------------------------------------------- shared maths */

/* |x| < RDELTA with x promoted to double (vfRayKernelCPU.cc:85, 206, 323). */
static inline bool tiny(float x) { return (double)fabsf(x) < RDELTA; }

/* The Cramer solve of vfRayKernelCPU.cc:73-122 (same as :380-429): plane
 * parameters p with [x y 1].p = target over the three face vertices. */
struct Planes { float z[3], s[3]; };

static inline void cramer(const F4& a, const F4& b, const F4& c, float tx, float ty, float tz,
                          float det, float out[3]) {
  float c0x = a.x, c0y = b.x, c0z = c.x; /* fmCol0 */
  float c1x = a.y, c1y = b.y, c1z = c.y; /* fmCol1 */
  float tmpx = ty - tz;
  float tmpy = (ty * c1z) - (c1y * tz);
  float tmpz = (c0y * tz) - (ty * c0z);
  out[0] = ((tx * (c1y - c1z)) - (c1x * tmpx) + tmpy) / det;
  out[1] = ((c0x * tmpx) - tx * (c0y - c0z) + tmpz) / det;
  out[2] = (c0x * (-tmpy) - (c1x * tmpz) + tx * (c0y * c1z - c1y * c0z)) / det;
}

static inline Planes face_planes(const F4& a, const F4& b, const F4& c) {
  float c0x = a.x, c0y = b.x, c0z = c.x;
  float c1x = a.y, c1y = b.y, c1z = c.y;
  float det = c0x * (c1y - c1z) - c1x * (c0y - c0z) + ((c0y * c1z) - (c1y * c0z));
  det = tiny(det) ? (float)RDELTA : det;
  Planes p;
  cramer(a, b, c, a.z, b.z, c.z, det, p.z);
  cramer(a, b, c, a.w, b.w, c.w, det, p.s);
  return p;
}

static inline float plane_eval(const float p[3], float ox, float oy) {
  return ox * p[0] + oy * p[1] + p[2];
}

/* obj = (pix*2)/(real)(res-1) - 1.0, vfRayKernelCPU.cc:250, 254. */
static inline float pix_to_obj(u32 pix, u32 res) {
  float q = (float)(pix * 2) / (float)(res - 1);
  return (float)((double)q - 1.0);
}

/* The exit-face test of vfRayKernelCPU.cc:312-350 over the rotated vertex
 * array Q (Q[k] = tet slot (entry+k)&3); candidate i = 1..3 is face
 * (entry+i)&3 with vertices Q[i], Q[i+1], Q[i+2] (mod 4).  Returns i or 0.
 * Also returns the barycentric pieces of the passing face. */
struct ExitHit { int i; float cy, cz, D; bool swapped; };

static inline bool tri_test(const F4& a, const F4& b, const F4& c, float ox, float oy,
                            float slack_is_rdelta, ExitHit* h) {
  (void)slack_is_rdelta;
  float v1x = b.x - a.x, v1y = b.y - a.y;
  float v2x = c.x - a.x, v2y = c.y - a.y;
  float cx = (v1x * v2y) - (v1y * v2x);
  if (tiny(cx)) return false;
  bool sw = false;
  if (cx < 0.0f) {
    float tx = v1x, ty = v1y;
    v1x = v2x; v1y = v2y; v2x = tx; v2y = ty;
    cx = -cx; sw = true;
  }
  float v0x = ox - a.x, v0y = oy - a.y;
  float cy = (v1x * v0y) - (v1y * v0x);
  float cz = (v0x * v2y) - (v0y * v2x);
  if ((cy >= 0.0f) && (cz >= 0.0f) && ((double)(cy + cz) <= (double)cx + RDELTA)) {
    h->cy = cy; h->cz = cz; h->D = cx; h->swapped = sw;
    return true;
  }
  return false;
}

static inline ExitHit find_exit(const F4 Q[4], float ox, float oy) {
  ExitHit h; h.i = 0; h.cy = h.cz = h.D = 0.0f; h.swapped = false;
  for (int i = 1; i < 4; ++i) {
    if (tri_test(Q[i], Q[(i + 1) & 3], Q[(i + 2) & 3], ox, oy, 1.0f, &h)) { h.i = i; return h; }
  }
  return h;
}

/* Barycentric interpolation over the triangle (a,b,c) from the inside-test
 * crosses: with e1 = b-a, e2 = c-a (swapped when their cross is negative),
 * cy = e1 x p, cz = p x e2, D = e1 x e2 > 0.  Weight of b is cz/D unswapped,
 * cy/D swapped; weight of c the other one.  1/D is rounded once and
 * multiplied in. */
static inline float bary(float A, float B, float C, const ExitHit& h) {
  float wb = h.swapped ? h.cy : h.cz;
  float wc = h.swapped ? h.cz : h.cy;
  float inv = 1.0f / h.D;
  return A + ((wb * (B - A) + wc * (C - A)) * inv);
}

/* ------------------------------------------------------------- psi (FP64) */

/* Psi(a, b) = int_0^1 exp(-a s - (b - a) s^2 / 2) ds with a = tau_front,
 * b = tau_back (the integral psiGammaTable512.h tabulates in gamma = tau/(1+tau)).
 * Composite 8-point Gauss-Legendre, panels scaled to the decay rate. */
static const double GLX[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                              -0.1834346424956498, 0.1834346424956498, 0.5255324099163290,
                              0.7966664774136267, 0.9602898564975363};
static const double GLW[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                              0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                              0.2223810344533745, 0.1012285362903763};

extern "C" double orc_psi_exact(double a, double b) {
  double m = a > b ? a : b;
  int panels = 1 + (int)(m * 0.5);
  if (panels > 4096) panels = 4096;
  double h = 1.0 / panels, sum = 0.0;
  for (int p = 0; p < panels; ++p) {
    double lo = p * h;
    for (int k = 0; k < 8; ++k) {
      double s = lo + 0.5 * h * (GLX[k] + 1.0);
      sum += GLW[k] * exp(-a * s - (b - a) * s * s * 0.5);
    }
  }
  return sum * 0.5 * h;
}

/* ---------------------------------------------------------- REF rendering */

struct RefOut {
  uint8_t* rgb8; float* rgba; double* rgba64; u32* entry_ext; u32* steps;
};

/* double -> u32 saturating (negative and NaN give 0), the conversion the
 * reference's GPU build applies to its linear colour sums. */
static inline u32 sat_u32(double v) { return v > 0.0 ? (v < 4294967296.0 ? (u32)v : 0xFFFFFFFFu) : 0u; }

/* One channel of the linear integration, vfRayKernel.cu:758-772 and
 * :787-792, in the text's float/double sub-expressions. */
static inline u32 lin_channel(u32 col, float cP, float cN, float aP, float aN, float tr, float dz) {
  double t1 = 0.5 * (double)(cP + cN) * (double)tr * (double)dz;
  double w = 3.0 * (double)cP * (double)aP + 5.0 * (double)cN * (double)aP + (double)(cP * aN) +
             3.0 * (double)cN * (double)aN;
  double t2 = 0.04166666 * w * (double)dz * (double)dz;
  u32 v = sat_u32((double)col + (t1 - t2));
  return v > 255u ? 255u : v;
}

/* One visible face of the reference's K3 (vfRayKernelCPU.cc:172-566), with
 * the face-plane memo (:234-245, :369-435) omitted: it is keyed by (tet, exit
 * face) and returns the values the same Cramer solve produces.  Returns 1 if
 * the face ended with the reference's face-level `return` (:365).
 * comp 0: PPI (:448-482); 1: linear integration (vfRayKernel.cu:738-806,
 * USEOPACITY unset: transparency 1, alpha weights 0, image = colour);
 * 2: linear with opacity (USEOPACITY: image = colour * opacity). */
static int ref_face(const float* V, const u32* T, const u32* con, const float* tf,
                    const float* psi, const u32* extf, u32 extIdx, const Planes& cf,
                    const u32 bb[4], u32 W, u32 H, float maxEdge, u32 comp, RefOut* o) {
  u32 t0 = extf[extIdx * 2 + 0], f0 = extf[extIdx * 2 + 1];
  F4 fv0 = ld4(V, T[t0 * 4 + ((f0 + 0) & 3)]);
  F4 fv1 = ld4(V, T[t0 * 4 + ((f0 + 1) & 3)]);
  F4 fv2 = ld4(V, T[t0 * 4 + ((f0 + 2) & 3)]);
  float v01x = fv1.x - fv0.x, v01y = fv1.y - fv0.y;
  float v02x = fv2.x - fv0.x, v02y = fv2.y - fv0.y;
  float cx = (v01x * v02y) - (v01y * v02x);
  cx = tiny(cx) ? (float)RDELTA : cx;
  if (cx < 0.0f) {
    float tx = v01x, ty = v01y;
    v01x = v02x; v01y = v02y; v02x = tx; v02y = ty;
    cx = -cx;
  }
  for (u32 px = bb[0]; px <= bb[2]; ++px) {
    float ox = pix_to_obj(px, W);
    for (u32 py = bb[1]; py <= bb[3]; ++py) {
      float oy = pix_to_obj(py, H);
      float v0px = ox - fv0.x, v0py = oy - fv0.y;
      float cy = (v01x * v0py) - (v01y * v0px);
      float cz = (v0px * v02y) - (v0py * v02x);
      if (!((cy >= 0.0f) && (cz >= 0.0f) && ((cy + cz) <= cx))) continue;
      float zPrev = plane_eval(cf.z, ox, oy);
      float sPrev = plane_eval(cf.s, ox, oy);
      if ((sPrev < 0.0f) || (sPrev > 1.0f)) break; /* :274 leaves the column */
      F4 cB = ld4(tf, (u32)(255.0f * sPrev));
      float ax = 0.0f, ay = 0.0f, az = 0.0f, Tr = 1.0f;
      double T64 = 1.0;
      double dB[4] = {cB.x, cB.y, cB.z, cB.w};
      double A64[3] = {0, 0, 0};
      const bool lin = comp != 0, opac = comp == 2;
      float cP[4] = {255.0f * cB.x, 255.0f * cB.y, 255.0f * cB.z, opac ? cB.w : 0.0f}; /* vfRayKernel.cu:463-468 */
      u32 col[3] = {0, 0, 0};
      float opacity = 0.0f;
      u32 tcur = t0, fent = f0, steps = 0;
      while (1) {
        if (++steps > MAX_STEPS) break;
        F4 fv[4];
        for (int k = 0; k < 4; ++k) fv[k] = ld4(V, T[tcur * 4 + k]);
        F4 Q[4];
        for (int k = 0; k < 4; ++k) Q[k] = fv[(fent + k) & 3];
        ExitHit h = find_exit(Q, ox, oy);
        if (h.i == 0) break; /* :353 */
        u32 fex = (fent + h.i) & 3;
        u32 tnext = con[tcur * 4 + fex];
        if (tnext != tcur) {
          u32 k;
          for (k = 0; k < 4; ++k) if (con[tnext * 4 + k] == tcur) break;
          if (k == 4) return 1; /* :365 */
          fent = k;
        }
        Planes nf = face_planes(fv[(fex + 0) & 3], fv[(fex + 1) & 3], fv[(fex + 2) & 3]);
        float zNext = plane_eval(nf.z, ox, oy);
        float sNext = plane_eval(nf.s, ox, oy);
        float dz = fabsf(zPrev - zNext);
        dz = ((double)dz < RDELTA) ? 0.0f : dz;
        if ((sNext < 0.0f) || (sNext > 1.0f)) break; /* :446 */
        if (lin) {
          /* vfRayKernel.cu:738-806; dz is not divided by maxEdgeLength */
          F4 tc = ld4(tf, (u32)(255.0f * sNext));
          float cN[4] = {255.0f * tc.x, 255.0f * tc.y, 255.0f * tc.z, opac ? tc.w : 0.0f};
          float tr = opac ? (float)(1.0 - (double)opacity) : 1.0f;
          u32 nc[3];
          for (int c = 0; c < 3; ++c) nc[c] = lin_channel(col[c], cP[c], cN[c], cP[3], cN[3], tr, dz);
          if (opac) opacity = (float)((double)opacity + (double)((cP[3] + cN[3]) * dz) / 2.0);
          for (int c = 0; c < 4; ++c) cP[c] = cN[c];
          for (int c = 0; c < 3; ++c) col[c] = nc[c];
          if (opac && opacity > 1.0f) { opacity = 1.0f; break; } /* :799-803 */
          if (tnext == tcur) break;
          tcur = tnext;
          zPrev = zNext;
          sPrev = sNext;
          continue;
        }
        dz /= maxEdge;
        F4 cF = ld4(tf, (u32)(255.0f * sNext));
        float taux = cF.w * dz, tauy = cB.w * dz;
        float zeta = (float)exp(-((double)taux * 0.5 + (double)tauy * 0.5));
        float gx = (float)((double)taux / (1.0 + (double)taux));
        float gy = (float)((double)tauy / (1.0 + (double)tauy));
        float tcx = (float)((double)gx + (0.5 / 512));
        float tcy = (float)((double)gy + (0.5 / 512));
        if ((tcx < 0.0f) || (tcx > 1.0f) || (tcy < 0.0f) || (tcy > 1.0f)) break; /* :468 */
        u32 ix = (u32)(tcx * 512.0f), iy = (u32)(tcy * 512.0f);
        if (ix > 511) ix = 511; /* texCoord == 1.0 would index row 512 */
        if (iy > 511) iy = 511;
        float ps = psi[ix * 512 + iy];
        float d = ps - zeta; /* float - float, :473 */
        float fx = (float)((double)cF.x * (1.0 - (double)ps) + (double)(cB.x * d));
        float fy = (float)((double)cF.y * (1.0 - (double)ps) + (double)(cB.y * d));
        float fz = (float)((double)cF.z * (1.0 - (double)ps) + (double)(cB.z * d));
        float fw = (float)(1.0 - (double)zeta);
        ax = (float)((double)fx + (1.0 - (double)fw) * (double)ax);
        ay = (float)((double)fy + (1.0 - (double)fw) * (double)ay);
        az = (float)((double)fz + (1.0 - (double)fw) * (double)az);
        Tr *= zeta;
        if (o->rgba64) {
          /* FP64 re-integration of the same segment: plane values in double
           * from the same face vertices, exact psi. */
          double dzz = fabs((double)zPrev - (double)zNext) / (double)maxEdge;
          double tf64 = (double)cF.w * dzz, tb64 = dB[3] * dzz;
          double z64 = exp(-(tf64 + tb64) * 0.5);
          double p64 = orc_psi_exact(tf64, tb64);
          double F[3] = {cF.x * (1.0 - p64) + dB[0] * (p64 - z64),
                         cF.y * (1.0 - p64) + dB[1] * (p64 - z64),
                         cF.z * (1.0 - p64) + dB[2] * (p64 - z64)};
          for (int c = 0; c < 3; ++c) A64[c] = F[c] + z64 * A64[c];
          T64 *= z64;
          dB[0] = cF.x; dB[1] = cF.y; dB[2] = cF.z; dB[3] = cF.w;
        }
        cB = cF;
        if (tnext == tcur) break; /* :535 */
        tcur = tnext;
        zPrev = zNext;
        sPrev = sNext;
      }
      if (ax > 1.0f) ax = 1.0f;
      if (ay > 1.0f) ay = 1.0f;
      if (az > 1.0f) az = 1.0f;
      u32 p = py * W + px;
      if (lin) {
        /* image byte: colour * opacity (vfRayKernel.cu:836-838) or colour */
        u32 b[3];
        for (int c = 0; c < 3; ++c) b[c] = opac ? (u32)((float)col[c] * opacity) : col[c];
        if (o->rgb8)
          for (int c = 0; c < 3; ++c) o->rgb8[p * 3 + c] = (uint8_t)b[c];
        if (o->rgba) {
          for (int c = 0; c < 3; ++c) o->rgba[p * 4 + c] = (float)b[c] / 255.0f;
          o->rgba[p * 4 + 3] = opac ? opacity : 1.0f;
        }
        if (o->entry_ext) o->entry_ext[p] = extIdx;
        if (o->steps) o->steps[p] = steps > MAX_STEPS ? MAX_STEPS : steps;
        continue;
      }
      if (o->rgb8) {
        o->rgb8[p * 3 + 0] = (uint8_t)(int)(255.0f * ax);
        o->rgb8[p * 3 + 1] = (uint8_t)(int)(255.0f * ay);
        o->rgb8[p * 3 + 2] = (uint8_t)(int)(255.0f * az);
      }
      if (o->rgba) {
        o->rgba[p * 4 + 0] = ax; o->rgba[p * 4 + 1] = ay; o->rgba[p * 4 + 2] = az;
        o->rgba[p * 4 + 3] = 1.0f - Tr;
      }
      if (o->rgba64) {
        for (int c = 0; c < 3; ++c) o->rgba64[p * 4 + c] = A64[c] > 1.0 ? 1.0 : A64[c];
        o->rgba64[p * 4 + 3] = 1.0 - T64;
      }
      if (o->entry_ext) o->entry_ext[p] = extIdx;
      if (o->steps) o->steps[p] = steps > MAX_STEPS ? MAX_STEPS : steps;
    }
  }
  return 0;
}

/* The reference's projectVisibleFaces bbox (vfRayKernelCPU.cc:141-167):
 * img = (uint)((res - 1.0) * (x + 1.0) / 2.0) in double. */
static inline u32 img_coord(float x, u32 res) {
  return (u32)(((double)res - 1.0) * ((double)x + 1.0) / 2.0);
}

/* main.cc:74-224 (executeKernel1, compaction with the duplicated last face,
 * executeKernel2, executeKernel3 in visible-face order).  V must be
 * normalized.  Returns the number of faces that ended with the face-level
 * `return` (vfRayKernelCPU.cc:365), or -1 on bad input.  out_nvis receives
 * numVisFaces including the duplicate. */
extern "C" long orc_render_ref(const float* V, u32 nV, const u32* T, u32 nT, const u32* con,
                               const u32* extf, u32 nExt, const float* tf, const float* psi,
                               const float* lmt, u32 W, u32 H, uint8_t* rgb8, float* rgba,
                               double* rgba64, u32* entry_ext, u32* steps, u32* out_nvis, u32 comp) {
  (void)nV; (void)nT;
  if (W < 2 || H < 2 || comp > 2) return -1;
  std::vector<Planes> pl(nExt);
  std::vector<u32> vis;
  for (u32 i = 0; i < nExt; ++i) {
    u32 t = extf[i * 2], f = extf[i * 2 + 1];
    F4 fv0 = ld4(V, T[t * 4 + ((f + 0) & 3)]);
    F4 fv1 = ld4(V, T[t * 4 + ((f + 1) & 3)]);
    F4 fv2 = ld4(V, T[t * 4 + ((f + 2) & 3)]);
    F4 op = ld4(V, T[t * 4 + ((f + 3) & 3)]);
    pl[i] = face_planes(fv0, fv1, fv2);
    float faceZ = op.x * pl[i].z[0] + op.y * pl[i].z[1] + pl[i].z[2];
    if (faceZ < op.z) pl[i].s[0] = -1.0f;            /* vfRayKernelCPU.cc:128 */
    if (pl[i].s[0] != -1.0f) vis.push_back(i);        /* main.cc:97, 112 */
  }
  if (vis.empty()) { if (out_nvis) *out_nvis = 0; return 0; }
  vis.push_back(vis.back());                          /* main.cc:124-126 */
  if (out_nvis) *out_nvis = (u32)vis.size();
  size_t np = (size_t)W * H;
  if (rgb8) memset(rgb8, 0, np * 3);
  if (rgba) memset(rgba, 0, np * 4 * sizeof(float));
  if (rgba64) memset(rgba64, 0, np * 4 * sizeof(double));
  if (entry_ext) for (size_t p = 0; p < np; ++p) entry_ext[p] = 0xFFFFFFFFu;
  if (steps) memset(steps, 0, np * sizeof(u32));
  RefOut o = {rgb8, rgba, rgba64, entry_ext, steps};
  long returns = 0;
  for (size_t k = 0; k < vis.size(); ++k) {
    u32 i = vis[k];
    u32 t = extf[i * 2], f = extf[i * 2 + 1];
    u32 bb[4] = {W - 1, H - 1, 0, 0};
    for (u32 j = 0; j < 3; ++j) {
      F4 fv = ld4(V, T[t * 4 + ((f + j) & 3)]);
      u32 ix = img_coord(fv.x, W), iy = img_coord(fv.y, H);
      if (ix < bb[0]) bb[0] = ix;
      if (iy < bb[1]) bb[1] = iy;
      if (ix > bb[2]) bb[2] = ix;
      if (iy > bb[3]) bb[3] = iy;
    }
    returns += ref_face(V, T, con, tf, psi, extf, i, pl[i], bb, W, H, lmt[0], comp, &o);
  }
  return returns;
}

/* ---------------------------------------------------------- FTB rendering */

/* View in the renderer's projective space P = (X, Y, Z, S):
 *   ortho:       X,Y,Z = rows of M applied to (x,y,z,1); S = s.
 *   perspective: xc,yc,dc = rows of M applied to (x,y,z,1);
 *                Z = 1/dc, X = xc*Z, Y = yc*Z, S = s*Z.
 * Planes map to planes, so rays through pixels are the lines X = ox, Y = oy
 * and march towards decreasing Z, as the reference's rays march -z.
 * Segment length in world units: perspective |1/Z1 - 1/Z0| *
 * sqrt(1 + (ox kx)^2 + (oy k)^2) with k = tan(fov_y/2), kx = k W/H; ortho
 * |Z1 - Z0| * k with k = world length of a unit step along -Z.  k and kx are
 * passed as `view_k` and `view_kx`. */
struct OView { u32 proj; float M[12]; float k; };

static inline F4 xform(const OView& v, const F4& p) {
  const float* M = v.M;
  float a = ((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3];
  float b = ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7];
  float c = ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11];
  F4 r;
  if (v.proj == 0) { r.x = a; r.y = b; r.z = c; r.w = p.w; }
  else { r.z = 1.0f / c; r.x = a * r.z; r.y = b * r.z; r.w = p.w * r.z; }
  return r;
}

struct D4 { double x, y, z, w; };
static inline D4 xform64(const OView& v, const F4& p) {
  const float* M = v.M;
  double a = (double)M[0] * p.x + (double)M[1] * p.y + (double)M[2] * p.z + (double)M[3];
  double b = (double)M[4] * p.x + (double)M[5] * p.y + (double)M[6] * p.z + (double)M[7];
  double c = (double)M[8] * p.x + (double)M[9] * p.y + (double)M[10] * p.z + (double)M[11];
  D4 r;
  if (v.proj == 0) { r.x = a; r.y = b; r.z = c; r.w = p.w; }
  else { r.x = a / c; r.y = b / c; r.z = 1.0 / c; r.w = p.w * r.z; }
  return r;
}

/* Order-preserving map of a float to u32 (larger float -> larger key). */
static inline u32 fkey(float f) {
  u32 u; memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* [0,1] clamp; NaN maps to 0. */
static inline float clamp01(float s) { return std::fmin(std::fmax(s, 0.0f), 1.0f); }

/* Barycentric value of a triangle at (ox,oy) in double. */
static inline double bary64(const D4& a, const D4& b, const D4& c, double ox, double oy,
                            double D4::*m) {
  double e1x = b.x - a.x, e1y = b.y - a.y, e2x = c.x - a.x, e2y = c.y - a.y;
  double px = ox - a.x, py = oy - a.y;
  double D = e1x * e2y - e1y * e2x;
  double u = (px * e2y - py * e2x) / D;   /* weight of b */
  double w = (e1x * py - e1y * px) / D;   /* weight of c */
  return a.*m + u * (b.*m - a.*m) + w * (c.*m - a.*m);
}

/* exp(x) for x <= 0 in binary32 (clamped at -87): Cody-Waite reduction
 * x = n ln2 + r with a two-part ln2, degree-7 Taylor polynomial on
 * |r| <= ln2/2 in Horner form with fused multiply-adds, scaled by 2^n.
 * The renderer's kernels evaluate the same operations, so the two agree bit
 * for bit; accuracy against libm is checked by the tests. */
extern "C" float orc_expf(float x) {
  x = std::max(x, -87.0f);
  float n = rintf(x * 1.44269504f);
  float r = fmaf(-n, 0.693145751953125f, x);
  r = fmaf(-n, 1.42860682030941723e-6f, r);
  float p = fmaf(r, 0.000198412701f, 0.00138888892f);
  p = fmaf(r, p, 0.00833333377f);
  p = fmaf(r, p, 0.0416666679f);
  p = fmaf(r, p, 0.166666672f);
  p = fmaf(r, p, 0.5f);
  p = fmaf(r, p, 1.0f);
  p = fmaf(r, p, 1.0f);
  return ldexpf(p, (int)n);
}

/* Vertex relative to the pixel (the renderer's walk works in these). */
static inline F4 rel(const F4& p, float ox, float oy) {
  F4 r; r.x = p.x - ox; r.y = p.y - oy; r.z = p.z; r.w = p.w;
  return r;
}
/* Edge function u -> v of the pixel with u, v relative to the pixel:
 * u.x*v.y - u.y*v.x.  The same edge walked v -> u gives exactly the negated
 * float, so the faces behind an entry face cover it without gaps. */
static inline float edge_rel(const F4& u, const F4& v) { return u.x * v.y - u.y * v.x; }
/* 1/x as the renderer's walk forms it: the seed 0x7EF311C3 - bits(x) and
 * three Newton steps y += y (1 - x y), each a pair of fused multiply-adds. */
extern "C" float orc_rcp(float x) {
  u32 u; memcpy(&u, &x, 4);
  u = 0x7EF311C3u - u;
  float y; memcpy(&y, &u, 4);
  for (int i = 0; i < 3; ++i) {
    float e = fmaf(-x, y, 1.0f);
    y = fmaf(y, e, y);
  }
  return y;
}

/* Depth value and scalar of the point with weights wA..wE (>= 0, D their
 * sum) over four vertices: ortho gives Z, perspective the depth 1/Z; both
 * from one reciprocal. */
static inline void face_values(bool persp, float D, float wA, const F4& A, float wB, const F4& B, float wC,
                               const F4& C, float wE, const F4& E, float* zd, float* s) {
  float zn = fmaf(wE, E.z, fmaf(wC, C.z, fmaf(wB, B.z, wA * A.z)));
  float sn = fmaf(wE, E.w, fmaf(wC, C.w, fmaf(wB, B.w, wA * A.w)));
  float rc = orc_rcp(persp ? zn : D);
  *zd = (persp ? D : zn) * rc;
  *s = clamp01(sn * rc);
}
static inline float pick3(bool c1, bool c2, float a, float b, float c) { return c1 ? a : (c2 ? b : c); }
static inline u32 tf_index(float s) {
  u32 i = (u32)(255.0f * s);
  return i > 255 ? 255 : i;
}

/* Front-to-back render of one view.  Entry candidates are every visible
 * external face whose projection holds the pixel (the reference's inclusive
 * inside test), keyed by (Z at the pixel, ext index) and walked front to
 * back; after an exit at depth Z_exit the walk resumes at the first entry with
 * Z < Z_exit.  Outputs rgba (W*H*4), rgba64 (FP64 re-integration, may be
 * NULL), entry_ext (front-most entry), steps.  Returns the number of
 * anomalous rays (no exit face found), or -1 on bad input. */
/* exp(x) as the walk forms exp(-p/2): on [-1/2, 0] the degree-6 Taylor
 * polynomial about -1/4 (coefficients e^(-1/4)/k! rounded to float) in
 * Horner form with fused multiply-adds; below -1/2, orc_expf. */
extern "C" float orc_exp_seg(float x) {
  if (x < -0.5f) return orc_expf(x);
  float u = x + 0.25f;
  float e = fmaf(u, 0.00108166772f, 0.00649000658f);
  e = fmaf(u, e, 0.0324500315f);
  e = fmaf(u, e, 0.129800126f);
  e = fmaf(u, e, 0.389400393f);
  e = fmaf(u, e, 0.778800786f);
  return fmaf(u, e, 0.778800786f);
}

/* psi(a, b) and zeta = exp(-(a + b)/2) of one segment, as the renderer
 * evaluates them (psiGammaTable512.h:12-19, vfRayKernelCPU.cc:458): with
 * p = (a + b)/2, q = (b - a)/2, E = exp(-p/2) (orc_exp_seg), psi = E K(p, q) and zeta = E E,
 * K(p, q) = int_{-1/2}^{1/2} exp(-p t + q (1/4 - t^2)) dt from its series in
 * p^(2j) q^k with coefficients M_jk / ((2j)! k!), M_jk = the integral of
 * t^(2j) (1/4 - t^2)^k over [-1/2, 1/2], kept to q-degree 5, 3, 1, 0 for
 * p^0, p^2, p^4, p^6, coefficients rounded to float, Horner in q then in p^2,
 * every step a fused multiply-add. */
extern "C" void orc_psi_zeta(float a, float b, float* psi, float* zeta) {
  float p = (a + b) * 0.5f, q = (b - a) * 0.5f;
  float P = p * p;
  float c0 = fmaf(q, 3.00625311e-06f, 6.61375670e-05f);
  c0 = fmaf(q, c0, 0.00119047624f);
  c0 = fmaf(q, c0, 0.0166666675f);
  c0 = fmaf(q, c0, 0.166666672f);
  c0 = fmaf(q, c0, 1.0f);
  float c1 = fmaf(q, 1.65343918e-05f, 0.000297619059f);
  c1 = fmaf(q, c1, 0.00416666688f);
  c1 = fmaf(q, c1, 0.0416666679f);
  float c2 = fmaf(q, 3.72023824e-05f, 0.000520833360f);
  float K = fmaf(P, 3.10019846e-06f, c2);
  K = fmaf(P, K, c1);
  K = fmaf(P, K, c0);
  float E = orc_exp_seg(p * -0.5f);
  *psi = E * K;
  *zeta = E * E;
}

extern "C" float orc_psi_series(float a, float b) {
  float ps, zeta;
  orc_psi_zeta(a, b, &ps, &zeta);
  return ps;
}

extern "C" long orc_render_ftb(const float* V, u32 nV, const u32* T, u32 nT, const u32* con,
                               const u32* extf, u32 nExt, const float* tf, const float* psi,
                               const float* lmt, u32 proj, const float* M, float view_k,
                               float view_kx, u32 W, u32 H, float term, float* rgba, double* rgba64,
                               u32* entry_ext, u32* steps_out, u32* n_entries_out, u32 composite) {
  (void)nT;
  if (W < 2 || H < 2) return -1;
  OView view; view.proj = proj; memcpy(view.M, M, sizeof(view.M)); view.k = view_k;
  std::vector<F4> P(nV);
  for (u32 i = 0; i < nV; ++i) P[i] = xform(view, ld4(V, i));
  /* visibility: the reference's K1 test in P space */
  std::vector<u32> vis;
  for (u32 i = 0; i < nExt; ++i) {
    u32 t = extf[i * 2], f = extf[i * 2 + 1];
    const F4& a = P[T[t * 4 + ((f + 0) & 3)]];
    const F4& b = P[T[t * 4 + ((f + 1) & 3)]];
    const F4& c = P[T[t * 4 + ((f + 2) & 3)]];
    const F4& op = P[T[t * 4 + ((f + 3) & 3)]];
    Planes pl = face_planes(a, b, c);
    float faceZ = op.x * pl.z[0] + op.y * pl.z[1] + pl.z[2];
    if (!(faceZ < op.z)) vis.push_back(i);
  }
  size_t np = (size_t)W * H;
  std::vector<u32> cnt(np + 1, 0);
  std::vector<u64> keys;
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<u32> fill;
    if (pass == 1) {
      for (size_t p = 0; p < np; ++p) cnt[p + 1] += cnt[p];
      keys.assign(cnt[np], 0);
      fill.assign(cnt.begin(), cnt.end() - 1);
    }
    for (u32 i : vis) {
      u32 t = extf[i * 2], f = extf[i * 2 + 1];
      const F4& a = P[T[t * 4 + ((f + 0) & 3)]];
      const F4& b = P[T[t * 4 + ((f + 1) & 3)]];
      const F4& c = P[T[t * 4 + ((f + 2) & 3)]];
      double lo[2] = {1e30, 1e30}, hi[2] = {-1e30, -1e30};
      const F4* vv[3] = {&a, &b, &c};
      for (int j = 0; j < 3; ++j) {
        double fx = ((double)W - 1.0) * ((double)vv[j]->x + 1.0) / 2.0;
        double fy = ((double)H - 1.0) * ((double)vv[j]->y + 1.0) / 2.0;
        lo[0] = std::min(lo[0], fx); hi[0] = std::max(hi[0], fx);
        lo[1] = std::min(lo[1], fy); hi[1] = std::max(hi[1], fy);
      }
      if (hi[0] < 0.0 || hi[1] < 0.0 || lo[0] > W - 1.0 || lo[1] > H - 1.0) continue;
      u32 bx0 = lo[0] < 0.0 ? 0 : (u32)lo[0], by0 = lo[1] < 0.0 ? 0 : (u32)lo[1];
      u32 bx1 = hi[0] > W - 1.0 ? W - 1 : (u32)hi[0], by1 = hi[1] > H - 1.0 ? H - 1 : (u32)hi[1];
      float v01x = b.x - a.x, v01y = b.y - a.y, v02x = c.x - a.x, v02y = c.y - a.y;
      float cx = (v01x * v02y) - (v01y * v02x);
      cx = tiny(cx) ? (float)RDELTA : cx;
      bool sw = false;
      if (cx < 0.0f) {
        float tx = v01x, ty = v01y;
        v01x = v02x; v01y = v02y; v02x = tx; v02y = ty;
        cx = -cx; sw = true;
      }
      for (u32 px = bx0; px <= bx1; ++px) {
        float ox = pix_to_obj(px, W);
        for (u32 py = by0; py <= by1; ++py) {
          float oy = pix_to_obj(py, H);
          float v0px = ox - a.x, v0py = oy - a.y;
          float cy = (v01x * v0py) - (v01y * v0px);
          float cz = (v0px * v02y) - (v0py * v02x);
          if (!((cy >= 0.0f) && (cz >= 0.0f) && ((cy + cz) <= cx))) continue;
          size_t p = (size_t)py * W + px;
          if (pass == 0) { cnt[p + 1]++; continue; }
          ExitHit h; h.cy = cy; h.cz = cz; h.D = cx; h.swapped = sw; h.i = 0;
          float Zp = bary(a.z, b.z, c.z, h);
          keys[fill[p]++] = ((u64)fkey(Zp) << 32) | i;
        }
      }
    }
  }
  const float maxEdge = lmt[0];
  const float inv_me = 1.0f / maxEdge;
  long anomalies = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : anomalies)
  for (long pl = 0; pl < (long)np; ++pl) {
    size_t p = (size_t)pl;
    u32 px = (u32)(p % W), py = (u32)(p / W);
    u64* K = keys.data() + cnt[p];
    u32 nk = cnt[p + 1] - cnt[p];
    std::sort(K, K + nk, [](u64 x, u64 y) { return x > y; });
    float out[4] = {0, 0, 0, 0};
    double out64[4] = {0, 0, 0, 0};
    u32 steps = 0;
    if (entry_ext) entry_ext[p] = nk ? (u32)(K[0] & 0xFFFFFFFFu) : 0xFFFFFFFFu;
    if (n_entries_out) n_entries_out[p] = nk;
    if (nk) {
      float ox = pix_to_obj(px, W), oy = pix_to_obj(py, H);
      float rayf = view_k;
      double rayf64 = 1.0;
      if (proj == 1) {
        const float ax = ox * view_kx, ay = oy * view_k;
        rayf = sqrtf((1.0f + ax * ax) + ay * ay);
        const double ax64 = (double)ox * view_kx, ay64 = (double)oy * view_k;
        rayf64 = sqrt(1.0 + ax64 * ax64 + ay64 * ay64);
      }
      float ar = 0, ag = 0, ab = 0, Tr = 1.0f;
      double ar64 = 0, ag64 = 0, ab64 = 0, T64 = 1.0;
      /* linear compositing: colour bytes and opacity (32-bit path), the same
       * integral without the per-step truncation in FP64 */
      const bool lin = composite != 0, opac = composite == 2;
      u32 lc[3] = {0, 0, 0};
      float lop = 0.0f;
      double lc64[3] = {0, 0, 0}, lop64 = 0.0;
      bool live32 = true, live64 = (rgba64 != nullptr);
      F4 cP = {0.0f, 0.0f, 0.0f, 0.0f};
      u32 cur = 0;
      const bool persp = proj == 1;
      const float rf = lin ? rayf : rayf * inv_me;
      while (cur < nk && (live32 || live64)) {
        u32 ei = (u32)(K[cur] & 0xFFFFFFFFu);
        u32 tcur = extf[ei * 2], fent = extf[ei * 2 + 1];
        /* entry face relative to the pixel, ordered to positive orientation */
        u32 i0 = T[tcur * 4 + ((fent + 0) & 3)], i1 = T[tcur * 4 + ((fent + 1) & 3)],
            i2 = T[tcur * 4 + ((fent + 2) & 3)], i3 = T[tcur * 4 + ((fent + 3) & 3)];
        F4 P0 = rel(P[i0], ox, oy), P1 = rel(P[i1], ox, oy), P2 = rel(P[i2], ox, oy);
        float o01 = edge_rel(P0, P1), o12 = edge_rel(P1, P2), o20 = edge_rel(P2, P0);
        if (o01 + o12 + o20 < 0.0f) {
          std::swap(P1, P2);
          std::swap(i1, i2);
          float n01 = -o20, n12 = -o12, n20 = -o01;
          o01 = n01; o12 = n12; o20 = n20;
        }
        float Dn = (o12 + o20) + o01;
        if (!(Dn > 0.0f)) {
          if (live32) anomalies++;
          break;
        }
        float prev, sprev;
        face_values(persp, Dn, o12, P0, o20, P1, o01, P2, 0.0f, P2, &prev, &sprev);
        if (live32) cP = ld4(tf, tf_index(sprev));
        /* FP64 entry */
        D4 a64 = xform64(view, ld4(V, i0));
        D4 b64 = xform64(view, ld4(V, i1));
        D4 c64 = xform64(view, ld4(V, i2));
        double Zp64 = bary64(a64, b64, c64, ox, oy, &D4::z);
        double Sp64 = bary64(a64, b64, c64, ox, oy, &D4::w);
        double sp64 = proj == 1 ? Sp64 / Zp64 : Sp64;
        sp64 = sp64 < 0 ? 0 : (sp64 > 1 ? 1 : sp64);
        F4 cP64 = ld4(tf, (u32)(255.0 * sp64));
        float exitZ = 0.0f;
        bool anomaly = false, anom32 = false;
        u32 iter = 0;
        while (1) {
          if (live32 && ++steps > MAX_STEPS) { anomaly = anom32 = true; break; }
          if (++iter > 4 * MAX_STEPS) { anomaly = true; break; }
          F4 P3 = rel(P[i3], ox, oy);
          float t0 = edge_rel(P3, P0), t1 = edge_rel(P3, P1), t2 = edge_rel(P3, P2);
          /* candidate faces (P3, Pi, Pj), (i,j) = (0,1), (1,2), (2,0): an exit
           * face when t_i >= 0, t_j <= 0 and not t_i = t_j = o_ij = 0 */
          bool p1 = t0 >= 0.0f && t1 <= 0.0f && !(t0 == 0.0f && t1 == 0.0f && o01 == 0.0f);
          bool p2 = t1 >= 0.0f && t2 <= 0.0f && !(t1 == 0.0f && t2 == 0.0f && o12 == 0.0f);
          bool p3 = t2 >= 0.0f && t0 <= 0.0f && !(t2 == 0.0f && t0 == 0.0f && o20 == 0.0f);
          if (!(p1 || p2 || p3)) { anomaly = true; anom32 = live32; break; }
          bool c1 = p1, c2 = !p1 && p2;
          if ((int)p1 + (int)p2 + (int)p3 > 1) {
            float inf = INFINITY;
            float z1 = p1 ? P3.z + (-t1 * (P0.z - P3.z) + t0 * (P1.z - P3.z)) / (t0 + o01 - t1) : inf;
            float z2 = p2 ? P3.z + (-t2 * (P1.z - P3.z) + t1 * (P2.z - P3.z)) / (t1 + o12 - t2) : inf;
            float z3 = p3 ? P3.z + (-t0 * (P2.z - P3.z) + t2 * (P0.z - P3.z)) / (t2 + o20 - t0) : inf;
            c1 = p1 && z1 <= z2 && z1 <= z3;
            c2 = !c1 && p2 && z2 <= z3;
          }
          /* exit face (P3, Pi, Pj): P3 weighs o_ij, Pi -t_j, Pj t_i, Pk 0 */
          float w3 = pick3(c1, c2, o01, o12, o20);
          float w0 = c1 ? -t1 : (c2 ? 0.0f : t2);
          float w1 = c1 ? t0 : (c2 ? -t2 : 0.0f);
          float w2 = c1 ? 0.0f : (c2 ? t1 : -t0);
          float Dx = ((w3 + w0) + w1) + w2;
          float zd, snext;
          face_values(persp, Dx, w3, P3, w0, P0, w1, P1, w2, P2, &zd, &snext);
          bool k0 = c2, k1 = !c1 && !c2, k2 = c1;
          u32 ib = c1 ? i0 : (c2 ? i1 : i2), ic = c1 ? i1 : (c2 ? i2 : i0);
          u32 ik = c1 ? i2 : (c2 ? i0 : i1);
          u32 sk = 0;
          while (sk < 3 && T[tcur * 4 + sk] != ik) ++sk;
          u32 fex = (sk + 1) & 3;
          if (live32 && lin) {
            /* vfRayKernel.cu:738-806 over dl object units, front to back */
            float dl = std::fmax(fabsf(zd - prev) * rf, 0.0f);
            prev = zd;
            F4 cN = ld4(tf, tf_index(snext));
            float pw = opac ? cP.w : 0.0f, nw = opac ? cN.w : 0.0f;
            float tr = opac ? (float)(1.0 - (double)lop) : 1.0f;
            const float pc[3] = {cP.x, cP.y, cP.z}, nc[3] = {cN.x, cN.y, cN.z};
            for (int ch = 0; ch < 3; ++ch) lc[ch] = lin_channel(lc[ch], 255.0f * pc[ch], 255.0f * nc[ch], pw, nw, tr, dl);
            if (opac) lop = (float)((double)lop + (double)((pw + nw) * dl) / 2.0);
            cP = cN;
            if (opac && lop > 1.0f) { lop = 1.0f; live32 = false; }
          } else if (live32) {
            float dl = std::fmax(fabsf(zd - prev) * rf, 0.0f);
            prev = zd;
            F4 cN = ld4(tf, tf_index(snext));
            float tf_ = cP.w * dl, tb_ = cN.w * dl;
            float ps, zeta;
            if (psi) {
              /* 512 * gamma for both ends with one division */
              float uf = 1.0f + tf_, ub = 1.0f + tb_;
              float q = 512.0f / (uf * ub);
              float gf = (tf_ * ub) * q, gb = (tb_ * uf) * q;
              u32 jf = (u32)(gf + 0.5f), jb = (u32)(gb + 0.5f);
              ps = psi[std::min(jf, 511u) * 512 + std::min(jb, 511u)];
              zeta = orc_expf(-(tf_ + tb_) * 0.5f);
            } else {
              orc_psi_zeta(tf_, tb_, &ps, &zeta);
            }
            float w1 = 1.0f - ps, w2 = ps - zeta;
            ar = fmaf(Tr, fmaf(cP.x, w1, cN.x * w2), ar);
            ag = fmaf(Tr, fmaf(cP.y, w1, cN.y * w2), ag);
            ab = fmaf(Tr, fmaf(cP.z, w1, cN.z * w2), ab);
            Tr = Tr * zeta;
            cP = cN;
            if (1.0f - Tr >= term) live32 = false;
          }
          if (live64) {
            D4 A64 = xform64(view, ld4(V, i3)), B64 = xform64(view, ld4(V, ib)),
               C64 = xform64(view, ld4(V, ic));
            double Zn64 = bary64(A64, B64, C64, ox, oy, &D4::z);
            double Sn64 = bary64(A64, B64, C64, ox, oy, &D4::w);
            double sn64 = proj == 1 ? Sn64 / Zn64 : Sn64;
            sn64 = sn64 < 0 ? 0 : (sn64 > 1 ? 1 : sn64);
            double dist = proj == 1 ? fabs(1.0 / Zn64 - 1.0 / Zp64) * rayf64 : fabs(Zp64 - Zn64) * (double)view_k;
            double dl = dist / (double)maxEdge;
            F4 cN = ld4(tf, (u32)(255.0 * sn64));
            if (lin) {
              /* colour c(u) linear between the samples times (1 - O(u)),
               * O(u) the linearly interpolated alpha integrated along the
               * segment: exact integral, clamped at 255 per step */
              double aP = opac ? cP64.w : 0.0, aN = opac ? cN.w : 0.0, tr = 1.0 - lop64;
              const double pc[3] = {cP64.x, cP64.y, cP64.z}, nc[3] = {cN.x, cN.y, cN.z};
              for (int ch = 0; ch < 3; ++ch) {
                double cp = 255.0 * pc[ch], cn = 255.0 * nc[ch];
                lc64[ch] = std::min(255.0, lc64[ch] + 0.5 * (cp + cn) * tr * dist -
                                               (3.0 * cp * aP + 5.0 * cn * aP + cp * aN + 3.0 * cn * aN) * dist * dist / 24.0);
              }
              if (opac) lop64 += (aP + aN) * dist / 2.0;
              cP64 = cN;
              Zp64 = Zn64;
              if (opac && lop64 > 1.0) { lop64 = 1.0; live64 = false; }
            } else {
              double tf_ = cP64.w * dl, tb_ = cN.w * dl;
              double zeta = exp(-(tf_ + tb_) * 0.5);
              double ps = orc_psi_exact(tf_, tb_);
              double w1 = 1.0 - ps, w2 = ps - zeta;
              ar64 += T64 * (cP64.x * w1 + cN.x * w2);
              ag64 += T64 * (cP64.y * w1 + cN.y * w2);
              ab64 += T64 * (cP64.z * w1 + cN.z * w2);
              T64 *= zeta;
              cP64 = cN;
              Zp64 = Zn64;
              if (1.0 - T64 >= (double)term) live64 = false;
            }
          }
          if (!live32 && !live64) break;
          u32 nb = con[tcur * 4 + fex];
          if (nb == tcur) { exitZ = persp ? 1.0f / zd : zd; break; }
          u32 k;
          for (k = 0; k < 4; ++k) if (con[nb * 4 + k] == tcur) break;
          if (k == 4) { anomaly = true; anom32 = false; break; }
          /* the exit face becomes the entry face: vertex k is replaced by P3 */
          float n01 = k0 ? t1 : (k1 ? -t0 : o01);
          float n12 = k1 ? t2 : (k2 ? -t1 : o12);
          float n20 = k2 ? t0 : (k0 ? -t2 : o20);
          if (k0) { P0 = P3; i0 = i3; }
          if (k1) { P1 = P3; i1 = i3; }
          if (k2) { P2 = P3; i2 = i3; }
          o01 = n01; o12 = n12; o20 = n20;
          i3 = T[nb * 4 + ((k + 3) & 3)];
          tcur = nb;
        }
        if (anom32) anomalies++;
        if (anomaly) break;
        if (!live32 && !live64) break;
        /* re-entry: first entry strictly behind the exit point */
        ++cur;
        while (cur < nk) {
          float z; u32 zk = (u32)(K[cur] >> 32);
          u32 u = (zk & 0x80000000u) ? (zk & 0x7FFFFFFFu) : ~zk;
          memcpy(&z, &u, 4);
          if (z < exitZ) break;
          ++cur;
        }
      }
      out[0] = ar; out[1] = ag; out[2] = ab; out[3] = 1.0f - Tr;
      out64[0] = ar64; out64[1] = ag64; out64[2] = ab64; out64[3] = 1.0 - T64;
      if (lin) {
        /* image byte: color * opacity (with opacity) or color, as byte / 255 */
        float op = opac ? lop : 1.0f;
        for (int ch = 0; ch < 3; ++ch) {
          u32 b = opac ? (u32)((float)lc[ch] * op) : lc[ch];
          out[ch] = (float)b / 255.0f;
          out64[ch] = (opac ? lc64[ch] * lop64 : lc64[ch]) / 255.0;
        }
        out[3] = op;
        out64[3] = opac ? lop64 : 1.0;
      }
    }
    if (rgba) for (int c = 0; c < 4; ++c) rgba[p * 4 + c] = out[c];
    if (rgba64) for (int c = 0; c < 4; ++c) rgba64[p * 4 + c] = out64[c];
    if (steps_out) steps_out[p] = steps > MAX_STEPS ? MAX_STEPS : steps;
  }
  return anomalies;
}
