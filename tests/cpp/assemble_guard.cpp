// Host-only check of launch_assemble's argument guard: a fused update needs
// a warm CG start and at least one CG iteration, and the launcher refuses
// the other combinations before it touches the device (so this runs without
// a GPU).  Prints one line per case: warm max_it fused -> hipError_t.
#include <cstdio>

#include "oc_internal.hpp"

int main() {
    oc::DevCloth c{};
    const int cases[][2] = {{0, 30}, {1, 0}, {0, 0}};
    for (const auto& k : cases) {
        oc::StepScalars s{};
        s.warm = k[0];
        s.max_it = k[1];
        const hipError_t e = oc::launch_assemble(c, s, nullptr, true);
        std::printf("%d %d 1 -> %d\n", k[0], k[1], (int)e);
    }
    std::printf("invalid %d\n", (int)hipErrorInvalidValue);
    return 0;
}
