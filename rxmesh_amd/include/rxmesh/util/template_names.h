// Inside a template, clang parses `query.dispatch<Op::FV>(...)` on an object
// of dependent type as a template call only when ordinary lookup of the name
// from the call site finds a class template (C++17 [basic.lookup.classref]/1);
// otherwise it asks for `query.template dispatch<...>`. The reference's
// users write the short form after `using namespace rxmesh;`. These
// declared-only class templates make that spelling parse unchanged; they are
// never defined or instantiated.
#pragma once

namespace rxmesh {
namespace template_names {
template <typename...>
struct dispatch;
template <typename...>
struct prologue;
template <typename...>
struct get_iterator;
template <typename...>
struct to_glm;
template <typename...>
struct from_glm;
}  // namespace template_names
using template_names::dispatch;
using template_names::from_glm;
using template_names::get_iterator;
using template_names::prologue;
using template_names::to_glm;
}  // namespace rxmesh
