/*
 * cleantf_main.cpp -- the `cleanTF old.off.tf new.tf` command (cleanTF.cc:
 * 14-78) on top of libvfr.so: converts a legacy transfer-function file to
 * the .tf text.  Exit status as the reference's: 1 for a usage error, 0
 * otherwise (its failures return false, i.e. 0, after a message).
 */
#include <cstdio>

#include "../../include/vfr.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "use: %s old.off.tf new.tf\nWhere *.tf is the transfer function file.\n", argv[0]);
    return 1;
  }
  printf("Reading %s . . . done!\nWriting %s . . . done!\n", argv[1], argv[2]);
  if (vfr_clean_tf(argv[1], argv[2]) != VFR_OK) fprintf(stderr, "%s\n", vfr_last_error());
  return 0;
}
