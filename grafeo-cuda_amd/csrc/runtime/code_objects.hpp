// Table of the gfx950 code objects linked into the library. The table itself
// is generated at build time (one entry per kernels/*.hip file, with the
// SHA-256 of the source and of the code object).
#pragma once

#include <cstdint>

namespace gc {

struct CodeObject {
  const char* name;             // kernel family, the .hip file stem
  const unsigned char* begin;   // offload bundle for gfx950
  const unsigned char* end;
  const char* source_sha256;    // of the .hip source (CRLF normalized to LF)
  const char* object_sha256;    // of the code object
};

}  // namespace gc

extern "C" const gc::CodeObject gc_embedded_code_objects[];
extern "C" const uint32_t gc_embedded_code_object_count;
