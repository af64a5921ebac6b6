// Handle types and helpers shared by the C entry points.
#pragma once

#include "engine.h"

#include <memory>

struct mrtEngine {
    std::unique_ptr<mrt::Engine> engine;
};

namespace mrt {

// Structs are accepted when non-NULL with structSize >= the size this
// library was built with.
template <typename T>
bool validStruct(const T* s) {
    return s && s->structSize >= sizeof(T);
}

inline Engine* unwrap(mrtEngine* e) { return e ? e->engine.get() : nullptr; }

} // namespace mrt
