// Host-only fuzz of the library's mesh-text parser (jst_mesh_parse), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_sanitized.py.  Seeded
// random texts: half of them near-valid meshes (4 x 8 or 4 x 9 points, odd number
// spellings, CRLF, blank lines, stray tokens), half token soup.  No GPU call is made.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../include/jst_gpu.h"

int main(int argc, char **argv)
{
    const int iters = argc > 1 ? atoi(argv[1]) : 20000;
    std::mt19937 rng(12345);
    const char *soup[] = {"1.5", "+2", "-0.0", "1e400", "-1e-400", "nan", "inf", "0x1p3", "1e", ".",
                          "1..2", " ", "\t", "\r", "\n", "\v", "\f", "abc", "12345678901234567890",
                          "3", "4", "8", "-", "(", "9e999999999"};
    const char *num[] = {"1.5", "+2", "-0.0", "1e400", "-1e-400", "nan", "inf", "3", ".25", "7E+2", "0.1"};
    const int nsoup = sizeof soup / sizeof *soup, nnum = sizeof num / sizeof *num;
    long ok = 0, rejected = 0;
    for (int it = 0; it < iters; ++it) {
        std::string s;
        if (it % 2) {
            const int n = rng() % 60;
            if (rng() % 2) s = "4 8\n";
            for (int k = 0; k < n; ++k) s += soup[rng() % nsoup];
        } else {
            s = (rng() % 4 ? "4 8" : " 4\t8 extra") + std::string(rng() % 2 ? "\n" : "\r\n");
            const int lines = 32 + (int)(rng() % 2);
            for (int k = 0; k < lines; ++k) {
                s += num[rng() % nnum];
                s += rng() % 2 ? " " : "\t";
                s += num[rng() % nnum];
                if (rng() % 50 == 0) s += soup[rng() % nsoup];
                s += rng() % 3 ? "\n" : "\r\n";
                if (rng() % 20 == 0) s += "\n";
            }
        }
        int a = 0, b = 0;
        if (jst_mesh_parse(s.c_str(), &a, &b, nullptr, 0) == JST_OK) {
            std::vector<double> xy(2 * (size_t)a * b);
            if (jst_mesh_parse(s.c_str(), &a, &b, xy.data(), xy.size()) != JST_OK) {
                fprintf(stderr, "second parse failed: %s\n", jst_last_error());
                return 1;
            }
            ++ok;
        } else {
            ++rejected;
            if (!jst_last_error()[0]) return 1;
        }
    }
    printf("mesh parse fuzz: %ld accepted, %ld rejected\n", ok, rejected);
    return ok > 0 && rejected > 0 ? 0 : 1;
}
