#pragma once
// Forward declarations of the public API (same set as the reference's include/helix/fwd.h).

namespace helix {

struct Bath;
struct ContextOptions;
struct Diagnostic;
class Context;
class Diagnostics;
struct HierarchySpec;
struct ReducedDensityShape;
struct RunResult;
struct SolverOptions;
struct SparseOperator;
struct System;
class HEOMSolver;

namespace examples {
System legacy_spin_glass_system();
}

}  // namespace helix
