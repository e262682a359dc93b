// ThreadSanitizer driver of the operator precompute (the host thread pool's parallel loops: Jacobi
// rounds, long-double products, M2L kernel spectra): builds the operators of several (P, images)
// repeatedly and checks that every repetition gives the same bits.  Prints "OK" on success.
#include <cstdint>
#include <cstdio>
#include <string>

#include "structure.h"

namespace {

uint64_t hash_ops(const fmm3dpp::Operators& op) {
    uint64_t h = 1469598103934665603ULL;
    auto add = [&](const std::vector<double>& v) {
        const unsigned char* q = reinterpret_cast<const unsigned char*>(v.data());
        for (size_t i = 0; i < v.size() * sizeof(double); i++) {
            h ^= q[i];
            h *= 1099511628211ULL;
        }
    };
    add(op.UT);
    add(op.VSinv);
    for (int k = 0; k < 8; k++) add(op.m2m[k]);
    add(op.gk8);
    add(op.gk27);
    return h;
}

}  // namespace

int main() {
    for (int P = 2; P <= 4; P++)
        for (int images = 0; images <= 2; images++) {
            uint64_t first = 0;
            for (int rep = 0; rep < 3; rep++) {
                fmm3dpp::Operators op;
                std::string err;
                if (fmm3dpp::build_operators(P, images, 1.0, op, err) != 0) {
                    printf("build_operators(P=%d, images=%d) failed: %s\n", P, images, err.c_str());
                    return 1;
                }
                const uint64_t h = hash_ops(op);
                if (rep == 0) first = h;
                else if (h != first) {
                    printf("P=%d images=%d: repetition %d differs\n", P, images, rep);
                    return 1;
                }
            }
        }
    printf("OK\n");
    return 0;
}
