// Internal declarations shared by the host C++ and the HIP translation units.
#pragma once

#include "../../include/dgk.h"

namespace dgk {

// Face f (0-based: 0 -y, 1 +x, 2 +y, 3 -x, 4 -z, 5 +z): normal axis, normal
// sign and opposite face (mod_mesh.f90:269-276, 366-394, 420-480).
struct FaceInfo {
  int axis;
  int sign;
  int opp;
};
extern const FaceInfo kFace[6];

int face_node_host(int Nq, int f, int a, int b);
long long neighbour_elem(int f, int i, int j, int k, int NeX, int NeY, int NeZ,
                         bool *wrapped);

// Records a message for dgk_last_error() and returns code.
int fail(int code, const char *fmt, ...);
void clear_error();

}  // namespace dgk
