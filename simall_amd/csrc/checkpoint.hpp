// checkpoint.hpp — library-internal access to the checkpoint's optional
// AMG trailer (the C ABI is in simall_amd.h).
#pragma once

#include "amg_setup.hpp"

struct mi_checkpoint;
struct mi_amg_host;

namespace mi
{

// Appends the hierarchy to an existing checkpoint file as a trailing block
// after the field section (tag "SIMAMGH1"), which io::Checkpoint::read does
// not reach and ignores.
int checkpoint_append_amg(const char* path, const AmgHierarchyHost& h);

// The hierarchy read from a checkpoint's trailer, or nullptr if it has none.
const AmgHierarchyHost* checkpoint_amg(const mi_checkpoint* c);

// The hierarchy behind an mi_amg_host handle (amg_host_api.cpp).
const AmgHierarchyHost* amg_host_hierarchy(const mi_amg_host* h);
// A new mi_amg_host handle holding H.
int amg_host_wrap(AmgHierarchyHost&& H, mi_amg_host** out);

} // namespace mi
