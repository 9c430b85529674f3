/*
 * vfray_main.cpp -- the `vfRay <volume>` command (vfRay.cu:737-936,
 * main.cc:242-380) on top of libvfr.so.
 *
 *   vfRay [--dir D] [--size W H] [--mode ref|ftb]
 *         [--composite ppi|linear|linear-opacity] [--rotate RX RY]... <volume>
 *
 * With no options it behaves as the reference: reads D<volume>.off with
 * D = "../tet_offs/", normalizes, then for .con, .tf, .extf and .lmt reads
 * the cache file or builds and writes it, in that order and with the
 * reference's messages.  A missing or unreadable .con is built from the
 * .incid file when that reads (the reference's buildCon over its lists), else
 * the .incid file is written and the connectivity built on the GPU by
 * sort-and-match; .extf is read whenever its count equals the .con's, also
 * after a .con rebuild, else built from .con.  Then it renders 512x512 along
 * -z and writes <volume>_gpu_img.ppm in the working directory.
 *
 * Each --rotate RX RY is one press of the CUDA viewer's space key with
 * rotation angles RX, RY degrees (vfRay.cu:617-640): the mesh is rotated by
 * Rx(RX) Ry(RY) as rotateVolume applies it and renormalized, after the cache
 * files are written; the limits stay those of the unrotated mesh, as in the
 * viewer.
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vfr.h"

static int fail(const char* what) {
  fprintf(stderr, "%s: %s\n", what, vfr_last_error());
  return 1;
}

int main(int argc, char** argv) {
  std::string dir = "../tet_offs/";
  uint32_t W = 512, H = 512, mode = VFR_MODE_REF, composite = VFR_COMPOSITE_PPI;
  std::vector<const char*> pos;
  std::vector<float> rot;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--dir") && i + 1 < argc) {
      dir = argv[++i];
      if (!dir.empty() && dir.back() != '/') dir += '/';
    } else if (!strcmp(argv[i], "--size") && i + 2 < argc) {
      W = (uint32_t)atoi(argv[++i]);
      H = (uint32_t)atoi(argv[++i]);
    } else if (!strcmp(argv[i], "--mode") && i + 1 < argc) {
      ++i;
      mode = !strcmp(argv[i], "ftb") ? VFR_MODE_FTB : VFR_MODE_REF;
    } else if (!strcmp(argv[i], "--rotate") && i + 2 < argc) {
      rot.push_back((float)atof(argv[++i]));
      rot.push_back((float)atof(argv[++i]));
    } else if (!strcmp(argv[i], "--composite") && i + 1 < argc) {
      /* the reference's USEPARTIALPREINT / USEOPACITY build switches
       * (vfRay.cu:38, vfRayKernel.cu:330) as an option */
      ++i;
      composite = !strcmp(argv[i], "linear")           ? VFR_COMPOSITE_LINEAR
                  : !strcmp(argv[i], "linear-opacity") ? VFR_COMPOSITE_LINEAR_OPACITY
                                                       : VFR_COMPOSITE_PPI;
    } else {
      pos.push_back(argv[i]);
    }
  }
  if (pos.size() != 1) {
    printf("use: %s model\nWhere model can be: spx, blunt, ...\n", argv[0]);
    return 1;
  }
  const std::string base = dir + pos[0];
  const std::string fnOff = base + ".off";
  printf("Reading %s file\n", fnOff.c_str());
  float* xyzs = nullptr;
  uint32_t *tets = nullptr, nV = 0, nT = 0;
  if (vfr_read_off(fnOff.c_str(), &xyzs, &nV, &tets, &nT) != VFR_OK) {
    printf("Cannot read %s aborting!\n", fnOff.c_str());
    return 1;
  }
  printf("Normalizing vertices\n");
  vfr_normalize_vertices(xyzs, nV);

  vfr_ctx* ctx = nullptr;
  if (vfr_ctx_create(0, &ctx) != VFR_OK) return fail("device");

  /* .con: read it, or build it from the .incid file (read, or built and
   * written first) as main.cc:271-299 does */
  uint32_t *con = nullptr, n_ext = 0;
  const std::string fnCon = base + ".con";
  printf("Reading %s file\n", fnCon.c_str());
  if (vfr_read_con(fnCon.c_str(), nT, &con, &n_ext) != VFR_OK) {
    printf("%s file not found!\n Computing conTet\n", fnCon.c_str());
    con = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)nT + 16);
    if (!con) return fail("memory");
    const std::string fnIncid = base + ".incid";
    printf("Reading %s file\n", fnIncid.c_str());
    uint32_t *off = nullptr, *ids = nullptr;
    if (vfr_read_incid(fnIncid.c_str(), nV, &off, &ids) == VFR_OK) {
      /* buildCon over the lists the file holds, on the host */
      int rc = vfr_con_from_incid(tets, nT, nV, off, ids, con, &n_ext);
      vfr_free(off);
      vfr_free(ids);
      if (rc != VFR_OK) return fail("con");
    } else {
      printf("%s file not found!\n Building incid\n", fnIncid.c_str());
      printf("Writing %s file\n", fnIncid.c_str());
      if (vfr_write_incid(fnIncid.c_str(), tets, nT, nV) != VFR_OK) return fail("incid");
      /* the same connectivity by sort-and-match on the device */
      vfr_mesh* tmp = nullptr;
      if (vfr_mesh_create(ctx, xyzs, nV, tets, nT, nullptr, nullptr, 0, nullptr, 0, &tmp) != VFR_OK)
        return fail("connectivity");
      vfr_mesh_info ti;
      vfr_mesh_get_info(tmp, &ti);
      n_ext = ti.n_ext_faces;
      int rc = vfr_mesh_connectivity(tmp, con, nullptr, nullptr);
      vfr_mesh_destroy(tmp);
      if (rc != VFR_OK) return fail("connectivity");
    }
    printf("Writing %s file\n", fnCon.c_str());
    if (vfr_write_con(fnCon.c_str(), con, nT, n_ext) != VFR_OK) return fail("con");
    printf("Deleting incidents in vertices\n");
  }
  /* .tf */
  float tf[1024];
  const std::string fnTF = base + ".tf";
  printf("Reading %s file\n", fnTF.c_str());
  if (vfr_read_tf(fnTF.c_str(), tf) != VFR_OK) {
    printf("%s file not found!\n Computing TF\n", fnTF.c_str());
    vfr_default_tf(tf);
    printf("Writing %s file\n", fnTF.c_str());
    if (vfr_write_tf(fnTF.c_str(), tf) != VFR_OK) return fail("tf");
  }
  /* .extf: read whenever it holds as many faces as .con says, also after a
   * .con rebuild (main.cc:312-328); otherwise buildExtF's scan of .con */
  uint32_t* extf = nullptr;
  const std::string fnExtF = base + ".extf";
  printf("Reading %s file\n", fnExtF.c_str());
  if (vfr_read_extf(fnExtF.c_str(), n_ext, &extf) != VFR_OK) {
    printf("%s file not found!\n Computing extFaces\n", fnExtF.c_str());
    std::vector<uint32_t> e;
    for (uint32_t t = 0; t < nT; ++t)
      for (uint32_t f = 0; f < 4; ++f)
        if (con[(size_t)t * 4 + f] == t) { e.push_back(t); e.push_back(f); }
    n_ext = (uint32_t)(e.size() / 2);
    extf = (uint32_t*)malloc(sizeof(uint32_t) * e.size() + 16);
    if (!extf) return fail("memory");
    if (!e.empty()) memcpy(extf, e.data(), sizeof(uint32_t) * e.size());
    printf("Writing %s file\n", fnExtF.c_str());
    if (vfr_write_extf(fnExtF.c_str(), extf, n_ext) != VFR_OK) return fail("extf");
  }
  /* .lmt */
  float lmt[3];
  const std::string fnLmt = base + ".lmt";
  printf("Reading %s file\n", fnLmt.c_str());
  if (vfr_read_lmt(fnLmt.c_str(), lmt) != VFR_OK) {
    printf("%s file not found!\n Computing limits\n", fnLmt.c_str());
    vfr_build_lmt(xyzs, nV, tets, nT, lmt);
    printf("Writing %s file\n", fnLmt.c_str());
    if (vfr_write_lmt(fnLmt.c_str(), lmt) != VFR_OK) return fail("lmt");
  }
  printf("Volume dataset information: %s\n", fnOff.c_str());
  printf("  # Tets: %d ; # Verts: %d\n", nT, nV);
  printf("  # ExtFaces: %d\n", n_ext);
  printf("  Max Edge Length: %f\n", lmt[0]);
  printf("  Max Z: %f ; Min Z: %f\n", lmt[1], lmt[2]);

  vfr_mesh* mesh = nullptr;
  if (vfr_mesh_create(ctx, xyzs, nV, tets, nT, con, extf, n_ext, lmt, VFR_MESH_MORTON, &mesh) != VFR_OK)
    return fail("mesh");
  vfr_mesh_info info;
  vfr_mesh_get_info(mesh, &info);
  printf("  Device preprocessing: %.2f ms\n", info.build_ms);

  for (size_t k = 0; k < rot.size(); k += 2) {
    float mv[16];
    vfr_gl_rotation(rot[k], rot[k + 1], mv);
    printf("* Rot X = %f ; Rot Y = %f\nNormalizing vertices\n", (double)rot[k], (double)rot[k + 1]);
    if (vfr_mesh_rotate(mesh, mv, VFR_MESH_NORMALIZE) != VFR_OK) return fail("rotate");
  }

  int tf_id = 0;
  if (vfr_tf_create(ctx, tf, &tf_id) != VFR_OK) return fail("tf upload");
  vfr_view v;
  memset(&v, 0, sizeof(v));
  v.width = W;
  v.height = H;
  v.projection = VFR_PROJ_ORTHO;
  v.tf_id = tf_id;
  v.ortho[0] = v.ortho[5] = v.ortho[10] = 1.0f;
  vfr_render_opts o;
  memset(&o, 0, sizeof(o));
  o.mode = mode;
  o.composite = composite;
  o.schedule = VFR_SCHED_BATCHED;
  std::vector<uint8_t> rgb((size_t)W * H * 3);
  vfr_frame_out fo;
  memset(&fo, 0, sizeof(fo));
  fo.rgb8 = rgb.data();
  vfr_stats st;
  if (vfr_render(ctx, mesh, &v, 1, &o, &fo, &st) != VFR_OK) return fail("render");
  printf("Visibility + entry lists: %.3f ms ; traversal: %.3f ms\n", st.ms_visibility, st.ms_traversal);
  printf("  # rays with an entry face: %llu ; tets visited: %llu\n",
         (unsigned long long)st.active_rays, (unsigned long long)st.steps);
  std::string fnImage = std::string(pos[0]) + "_gpu_img.ppm";
  printf("Saving ppm image: %s\n", fnImage.c_str());
  if (vfr_write_ppm(fnImage.c_str(), rgb.data(), W, H) != VFR_OK) return fail("ppm");
  printf("Clean up device\n");
  vfr_mesh_destroy(mesh);
  vfr_ctx_destroy(ctx);
  vfr_free(xyzs);
  vfr_free(tets);
  vfr_free(con);
  vfr_free(extf);
  return 0;
}
