#include <stdexcept>

#include "mixed.hpp"

namespace exa {

double mixed_loglik(Context&, KernelId, const double*, int, const MixedOptions&, double*, double*, int*) {
  throw std::runtime_error("mixed precision path not built yet");
}

}  // namespace exa
