// Host build of libtp's FP64 special functions (csrc/tp_host_math.hpp) for the
// CPU test suite.  Reads commands on stdin, prints results with 17 digits.
//   w l1 m1 l2 m2 l3 m3   -> wigner3j
//   p l m theta           -> sphPlm
//   s l m LMAX            -> seanindex
//   g n                   -> n Gauss-Legendre nodes and weights
#include <cstdio>
#include <vector>
#include "tp_host_math.hpp"

int main()
{
    char c;
    while (std::scanf(" %c", &c) == 1) {
        if (c == 'w') {
            int a[6];
            if (std::scanf("%d %d %d %d %d %d", a, a + 1, a + 2, a + 3, a + 4, a + 5) != 6) return 1;
            std::printf("%.17g\n", tph::wigner3j(a[0], a[1], a[2], a[3], a[4], a[5]));
        } else if (c == 'p') {
            int l, m;
            double th;
            if (std::scanf("%d %d %lf", &l, &m, &th) != 3) return 1;
            std::printf("%.17g\n", tph::sphPlm(l, m, th));
        } else if (c == 's') {
            int l, m, L;
            if (std::scanf("%d %d %d", &l, &m, &L) != 3) return 1;
            std::printf("%lld\n", (long long)tph::seanindex(l, m, L));
        } else if (c == 'g') {
            int n;
            if (std::scanf("%d", &n) != 1) return 1;
            std::vector<double> x, w;
            tph::gauss_legendre(n, x, w);
            for (int i = 0; i < n; ++i) std::printf("%.17g %.17g\n", x[i], w[i]);
        }
    }
    return 0;
}
