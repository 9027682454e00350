// Error state shared by the C ABI entry points.
#pragma once
#include <string>

#include "status.hpp"

namespace vfvm {
struct Mesh;
}

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
// Clears the calling thread's last error and returns VFVM_OK.
int clear_error();
// Mesh behind an opaque handle (defined with the mesh entry points).
const vfvm::Mesh* vfvm_mesh_ptr(const struct vfvm_mesh* mesh);
