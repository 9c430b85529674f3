/*
 * ref_harness.cpp -- drives the reference CPU path (vfRayPreComp.cc and
 * vfRayKernelCPU.cc, compiled from the reference tree into oracle/_ref/) to
 * record what main.cc does not print: which visible face wrote each pixel.
 * TEST INFRASTRUCTURE: used only to pin the oracle.
 *
 *   ref_harness <dataset prefix> <width> <height> <out prefix> [rotations]
 *
 * Runs main.cc's sequence (main.cc:257-342: readOff, normalizeVertices, the
 * .con/.incid, .tf, .extf, .lmt cache-or-build steps), K1 over the external
 * faces, main.cc's compaction with its duplicated last face (main.cc:93-135),
 * K2, then K3 per visible face in order (main.cc:209-218) into the image.
 * Each face is also rendered alone into two images pre-filled with 0x00 and
 * 0xFF; a pixel that holds the same value in both was written by that face,
 * so the last face that writes a pixel is its owner.
 *
 * [rotations]: a file of k 4x4 float32 matrices; after the cache steps each
 * is applied as the CUDA viewer's space key does (vfRay.cu:617-640):
 * rotateVolume(matrix), normalizeVertices(), limits kept.
 *
 * Writes <out>.rgb (width*height*3), <out>.owner (uint32 ext-face index of
 * the owning face, 0xFFFFFFFF if none), <out>.verts (the normalized vertex
 * list, float32) and <out>.txt (counts).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vfRayPreComp.h"
#include "vfRayKernelCPU.h"

uint numVerts, numTets;
real* h_vertList;
uint* h_tetList;
uint* h_extFaces;
uint* h_conTet;
real* h_transferFunction;
uint numExtFaces;
real h_maxEdgeLength, h_maxZ, h_minZ;

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    fprintf(stderr, "use: %s <dataset prefix> <width> <height> <out prefix> [rotations]\n", argv[0]);
    return 1;
  }
  std::string ds = argv[1];
  uint W = (uint)atoi(argv[2]), H = (uint)atoi(argv[3]);
  std::string out = argv[4];
  if (!readOff((ds + ".off").c_str())) { fprintf(stderr, "cannot read off\n"); return 1; }
  normalizeVertices();
  if (!readCon((ds + ".con").c_str())) {
    if (!readIncid((ds + ".incid").c_str())) {
      buildIncid();
      writeIncid((ds + ".incid").c_str());
    }
    buildCon();
    writeCon((ds + ".con").c_str());
    deleteIncid();
  }
  if (!readTF((ds + ".tf").c_str())) {
    buildTF();
    writeTF((ds + ".tf").c_str());
  }
  if (!readExtF((ds + ".extf").c_str())) {
    buildExtF();
    writeExtF((ds + ".extf").c_str());
  }
  if (!readLmt((ds + ".lmt").c_str())) {
    buildLmt();
    writeLmt((ds + ".lmt").c_str());
  }
  if (argc == 6) {
    FILE* fr = fopen(argv[5], "rb");
    if (!fr) { fprintf(stderr, "cannot read rotations\n"); return 1; }
    real mv[16];
    while (fread(mv, sizeof(real), 16, fr) == 16) {
      rotateVolume(mv);
      normalizeVertices();
    }
    fclose(fr);
  }
  std::vector<faceData> vis(numExtFaces);
  for (uint i = 0; i < numExtFaces; ++i)
    findVisibleFaces(vis.data(), h_vertList, h_tetList, h_extFaces, i);
  std::vector<faceData> vf;
  std::vector<uint> ef, idx;
  for (uint i = 0; i < numExtFaces; ++i)
    if (vis[i].sParams.x != -1.0) {
      vf.push_back(vis[i]);
      ef.push_back(h_extFaces[i * 2]);
      ef.push_back(h_extFaces[i * 2 + 1]);
      idx.push_back(i);
    }
  if (vf.empty()) { fprintf(stderr, "no visible face\n"); return 1; }
  vf.push_back(vf.back());
  ef.push_back(ef[ef.size() - 2]);
  ef.push_back(ef[ef.size() - 2]);
  idx.push_back(idx.back());
  uint nvis = (uint)vf.size();
  std::vector<projectedFace> pf(nvis);
  for (uint i = 0; i < nvis; ++i)
    projectVisibleFaces(pf.data(), h_vertList, h_tetList, ef.data(), vf.data(), i, W, H);
  std::vector<uchar> img((size_t)W * H * 3, 0), a((size_t)W * H * 3), b((size_t)W * H * 3);
  std::vector<uint> owner((size_t)W * H, 0xFFFFFFFFu);
  for (uint i = 0; i < nvis; ++i) {
    vfRay(img.data(), h_vertList, h_tetList, ef.data(), vf.data(), pf.data(), h_conTet,
          h_transferFunction, i, W, H, h_maxEdgeLength, h_maxZ, h_minZ);
    const uint4& bb = pf[i].bBox;
    for (uint y = bb.y; y <= bb.w && y < H; ++y)
      for (uint x = bb.x; x <= bb.z && x < W; ++x) {
        size_t p = ((size_t)y * W + x) * 3;
        a[p] = a[p + 1] = a[p + 2] = 0x00;
        b[p] = b[p + 1] = b[p + 2] = 0xFF;
      }
    vfRay(a.data(), h_vertList, h_tetList, ef.data(), vf.data(), pf.data(), h_conTet,
          h_transferFunction, i, W, H, h_maxEdgeLength, h_maxZ, h_minZ);
    vfRay(b.data(), h_vertList, h_tetList, ef.data(), vf.data(), pf.data(), h_conTet,
          h_transferFunction, i, W, H, h_maxEdgeLength, h_maxZ, h_minZ);
    for (uint y = bb.y; y <= bb.w && y < H; ++y)
      for (uint x = bb.x; x <= bb.z && x < W; ++x) {
        size_t p = (size_t)y * W + x;
        if (a[p * 3] == b[p * 3]) owner[p] = idx[i];
      }
  }
  FILE* f = fopen((out + ".rgb").c_str(), "wb");
  if (!f) return 1;
  fwrite(img.data(), 1, img.size(), f);
  fclose(f);
  f = fopen((out + ".owner").c_str(), "wb");
  if (!f) return 1;
  fwrite(owner.data(), 4, owner.size(), f);
  fclose(f);
  f = fopen((out + ".verts").c_str(), "wb");
  if (!f) return 1;
  fwrite(h_vertList, sizeof(real), (size_t)numVerts * 4, f);
  fclose(f);
  f = fopen((out + ".txt").c_str(), "w");
  if (!f) return 1;
  fprintf(f, "numVerts %u\nnumTets %u\nnumExtFaces %u\nnumVisFaces %u\n", numVerts, numTets,
          numExtFaces, nvis);
  fclose(f);
  return 0;
}
