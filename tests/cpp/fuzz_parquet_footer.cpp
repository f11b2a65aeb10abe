// Footer fuzzing of gbagg_parquet_open / _shape / _column / _row_group_rows
// (no GPU needed): random byte changes and cuts inside a file's footer, each
// image in a heap block of its exact size, so that a build with host-side
// AddressSanitizer reports any read outside it. Build and run (repo root):
//   hipcc --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -shared \
//     -Xclang -target-feature -Xclang -packed-fp32-ops -Xarch_host -fsanitize=address \
//     -Iinclude -o /tmp/libslothdb_gpu_asan.so slothdb_amd/csrc/gbagg.hip
//   hipcc -O1 -g -std=c++17 -fsanitize=address -Iinclude -o /tmp/fuzz_footer \
//     tests/cpp/fuzz_parquet_footer.cpp -L/tmp -lslothdb_gpu_asan -Wl,-rpath,/tmp
//   ASAN_OPTIONS=detect_leaks=0 /tmp/fuzz_footer some.parquet 200000
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <vector>
#include <random>
#include "slothdb_gpu.h"
int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s file.parquet [trials]\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    const int trials = argc > 2 ? atoi(argv[2]) : 200000;
    std::vector<uint8_t> img(64 << 20);
    size_t n = fread(img.data(), 1, img.size(), f);
    fclose(f);
    img.resize(n);
    if (n < 12) return 2;
    uint32_t flen; memcpy(&flen, &img[n - 8], 4);
    size_t f0 = n - 8 - flen;
    std::mt19937_64 rng(5);
    int ok = 0, bad = 0;
    for (int trial = 0; trial < trials; trial++) {
        std::vector<uint8_t> b = img;
        int m = 1 + rng() % 4;
        for (int i = 0; i < m; i++) { size_t pos = f0 + rng() % flen; b[pos] = (rng() & 1) ? (uint8_t)rng() : b[pos] ^ (1u << (rng() % 8)); }
        if (trial % 7 == 0) {
            uint32_t cut = 1 + rng() % (flen - 1);
            std::vector<uint8_t> c(b.begin(), b.begin() + (f0 + flen - cut));
            c.insert(c.end(), b.end() - 8, b.end());
            uint32_t nl = flen - cut; memcpy(&c[c.size() - 8], &nl, 4);
            b.swap(c);
        }
        // exact-size heap copy so ASan sees reads past the end
        uint8_t *h = new uint8_t[b.size()];
        memcpy(h, b.data(), b.size());
        gbagg_parquet *p = nullptr;
        if (gbagg_parquet_open(h, b.size(), &p) == GBAGG_OK) {
            uint32_t nrg, nc; uint64_t rows;
            gbagg_parquet_shape(p, &nrg, &nc, &rows);
            for (uint32_t c = 0; c < nc; c++) { gbagg_type t; uint8_t nl; const char *name; gbagg_parquet_column(p, c, &t, &nl, &name); if (name) (void)strlen(name); }
            for (uint32_t r = 0; r < nrg; r++) { uint64_t x; gbagg_parquet_row_group_rows(p, r, &x); }
            gbagg_parquet_close(p);
            ok++;
        } else bad++;
        delete[] h;
    }
    printf("opened %d rejected %d\n", ok, bad);
}
