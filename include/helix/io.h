#pragma once
// Added to the reference's public API: trajectory output of an ObservableTrace run in the legacy CLI's
// outputEnergy.txt format (src/cli_compatibility.cu:80-98,200-238).  Checkpoints are members of
// helix::Context (save_checkpoint / load_checkpoint); the file format is documented in helix_io.h.

#include <cstdint>
#include <iosfwd>
#include <string>

namespace helix {

struct RunResult;

// How the energy column is printed: Float32 rounds it to single precision first, as the compiled
// reference does (its Complex is single precision); Float64 prints the FP64 value.
enum class EnergyValues : std::uint8_t { Float32, Float64 };

// One "<time> <energy>" row per entry of result.times, with 8 significant digits, energy from
// result.energy_diagonal (sum_i Re H_ii Re rho_ii).
void write_energy_trace(std::ostream& out, const RunResult& result, EnergyValues values = EnergyValues::Float32);
bool write_energy_trace(const std::string& path, const RunResult& result, EnergyValues values = EnergyValues::Float32);

}  // namespace helix
