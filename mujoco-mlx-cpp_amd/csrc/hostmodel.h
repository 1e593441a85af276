// Flattening of a compiled model into the kernel tables (one blob, offsets
// resolved against a host or device base address).
#pragma once
#include <string>
#include <vector>

#include "dmodel.h"
#include "model.h"

namespace mjh {

struct Blob {
  std::vector<char> bytes;
  // offsets of every table, in bytes
  std::vector<std::pair<std::string, size_t>> offs;
  mjk::DModel proto;  // sizes and scalars; pointers are offsets until bind()
  mjk::DModel bind(const char* base) const;
};

// Builds the kernel tables. Throws std::runtime_error when the model needs a
// feature the kernels do not implement (the message names it).
Blob build_blob(const Model& m, std::vector<CollisionPair>* pairs_out);

// upper bound of contacts one env can produce (sum over pairs of the narrow
// phase's maximum) and of constraint rows
int max_contacts(const Model& m, const std::vector<CollisionPair>& pairs);
int max_rows(const Model& m, const std::vector<CollisionPair>& pairs);

// Per-env capacities, the reference's (src/batched.cpp:290 MAX_CONTACTS_PER_ENV
// = 128; src/compat/mkx_kernels/solver.hpp:28-33 row budget 72 for nv <= 32,
// else 256, MJMLX_MAX_EFC overriding it). Buffers are sized to what the model
// can produce under those budgets: contacts beyond the ones the row budget
// can admit are counted, not stored. maxcon_override / efc_cap_override > 0
// replace the contact cap and the row budget.
struct Caps { int maxcon, maxefc, efc_cap, con_cap; };
Caps default_caps(const Model& m, const std::vector<CollisionPair>& pairs, int maxcon_override = 0,
                  int efc_cap_override = 0);

// workspace field table for debug dumps: name -> (offset in floats, length, is_int)
struct LayoutField { const char* name; int off; int len; int kind; };  // kind 0 float, 1 int, 2 double
std::vector<LayoutField> layout_fields(const Model& m, const mjk::Layout& L);

// raw model field by name (integers as doubles); -1 for an unknown name
int model_field(const Model& m, const char* name, double* out, int cap);

}  // namespace mjh
