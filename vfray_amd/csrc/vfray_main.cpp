/*
 * vfray_main.cpp -- the `vfRay <volume>` command (vfRay.cu:737-936,
 * main.cc:242-380) on top of libvfr.so.
 *
 *   vfRay [--dir D] [--size W H] [--mode ref|ftb]
 *         [--composite ppi|linear|linear-opacity] [--rotate RX RY]... <volume>
 *
 * With no options it behaves as the reference: reads D<volume>.off with
 * D = "../tet_offs/", normalizes, then for .con (.incid on the way), .tf,
 * .extf and .lmt reads the cache file or builds and writes it, in that order
 * (.extf's count is checked against .con's header); renders 512x512 along -z
 * and writes <volume>_gpu_img.ppm in the working directory.  Connectivity and
 * external faces are built on the GPU.
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

  /* .con (and .incid) */
  uint32_t *con = nullptr, n_ext = 0;
  const std::string fnCon = base + ".con";
  printf("Reading %s file\n", fnCon.c_str());
  bool con_built = vfr_read_con(fnCon.c_str(), nT, &con, &n_ext) != VFR_OK;
  vfr_mesh* mesh = nullptr;
  if (con_built) {
    printf("%s file not found!\n Computing conTet\n", fnCon.c_str());
    const std::string fnIncid = base + ".incid";
    FILE* fi = fopen(fnIncid.c_str(), "r");
    if (fi) {
      fclose(fi);
    } else {
      printf("%s file not found!\n Building incid\n", fnIncid.c_str());
      printf("Writing %s file\n", fnIncid.c_str());
      if (vfr_write_incid(fnIncid.c_str(), tets, nT, nV) != VFR_OK) return fail("incid");
    }
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
  /* .extf: needs numExtFaces from .con; build on the device if absent */
  uint32_t* extf = nullptr;
  const std::string fnExtF = base + ".extf";
  bool extf_built = true;
  if (!con_built) {
    printf("Reading %s file\n", fnExtF.c_str());
    extf_built = vfr_read_extf(fnExtF.c_str(), n_ext, &extf) != VFR_OK;
  }
  /* .lmt */
  float lmt[3];
  const std::string fnLmt = base + ".lmt";
  bool lmt_built = vfr_read_lmt(fnLmt.c_str(), lmt) != VFR_OK;
  if (lmt_built) vfr_build_lmt(xyzs, nV, tets, nT, lmt);

  if (vfr_mesh_create(ctx, xyzs, nV, tets, nT, con, extf_built ? nullptr : extf, n_ext, lmt,
                      VFR_MESH_MORTON, &mesh) != VFR_OK)
    return fail("mesh");
  vfr_mesh_info info;
  vfr_mesh_get_info(mesh, &info);
  if (con_built) {
    std::vector<uint32_t> c((size_t)nT * 4);
    vfr_mesh_connectivity(mesh, c.data(), nullptr, nullptr);
    printf("Writing %s file\n", fnCon.c_str());
    if (vfr_write_con(fnCon.c_str(), c.data(), nT, info.n_ext_faces) != VFR_OK) return fail("con");
    printf("Deleting incidents in vertices\n");
    printf("Reading %s file\n", fnExtF.c_str());
  }
  if (extf_built) {
    printf("%s file not found!\n Computing extFaces\n", fnExtF.c_str());
    std::vector<uint32_t> e((size_t)info.n_ext_faces * 2);
    vfr_mesh_connectivity(mesh, nullptr, e.data(), nullptr);
    printf("Writing %s file\n", fnExtF.c_str());
    if (vfr_write_extf(fnExtF.c_str(), e.data(), info.n_ext_faces) != VFR_OK) return fail("extf");
  }
  printf("Reading %s file\n", fnLmt.c_str());
  if (lmt_built) {
    printf("%s file not found!\n Computing limits\n", fnLmt.c_str());
    printf("Writing %s file\n", fnLmt.c_str());
    if (vfr_write_lmt(fnLmt.c_str(), lmt) != VFR_OK) return fail("lmt");
  }
  printf("Volume dataset information: %s\n", fnOff.c_str());
  printf("  # Tets: %d ; # Verts: %d\n", nT, nV);
  printf("  # ExtFaces: %d\n", info.n_ext_faces);
  printf("  Max Edge Length: %f\n", lmt[0]);
  printf("  Max Z: %f ; Min Z: %f\n", lmt[1], lmt[2]);
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
