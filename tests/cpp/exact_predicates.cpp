// Host build of the library's orientation predicates (svm_exact.h, the same
// header the GPU kernels use): reads lines "a0 a1 a2 b0 b1 b2 c0 c1 c2 [d0 d1
// d2]" of doubles in hexadecimal float notation and prints, per line, the
// sign of orient_origin_sign (9 numbers) or orient3d_sign (12 numbers).
#include <cstdio>
#include <cstdlib>
#include "../../curvedspacevertexmodel_amd/csrc/svm_exact.h"

int main() {
    char line[4096];
    while (std::fgets(line, sizeof line, stdin)) {
        double v[12];
        int n = 0;
        char* p = line;
        while (n < 12) {
            char* end = nullptr;
            const double x = std::strtod(p, &end);
            if (end == p) break;
            v[n++] = x;
            p = end;
        }
        if (n == 9) std::printf("%d\n", svm::orient_origin_sign(v, v + 3, v + 6));
        else if (n == 12) std::printf("%d\n", svm::orient3d_sign(v, v + 3, v + 6, v + 9));
        else return 2;
    }
    return 0;
}
