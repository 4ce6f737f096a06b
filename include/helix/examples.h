#pragma once

#include <helix/types.h>

namespace helix::examples {

// The reference's compiled legacy spin glass (N = 1024, one Drude-Lorentz bath, src/parameters.cu:11-39).
System legacy_spin_glass_system();

// added: the seeded transverse-field spin glass with one sigma^z coupling per spin,
// H = sum_{i<j} J_ij sz_i sz_j + h sum_i sx_i, J_ij ~ N(0,1) from SplitMix64(seed).
System spin_glass_system(int spins, unsigned long long seed, double h = 1.0);

}  // namespace helix::examples
