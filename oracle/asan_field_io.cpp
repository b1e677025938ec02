// Sanitizer driver of the grid text I/O (ppafm_amd/csrc/field_io.cpp; test
// infrastructure): writes XSF data blocks of a seeded grid (contiguous and
// strided views, 1 and 4 threads) and reads them back with 1 and 3 threads,
// in buffers of exactly the documented sizes.  Built and run by
// `make -C oracle asan`.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

extern "C" int ppafm_xsf_write_data(const char *fname, int append, const double *data, int nx, int ny, int nz,
                                    long long sx, long long sy, long long sz, int nthreads);
extern "C" int ppafm_read_numbers(const char *fname, int noline, long long count, double *out, int nthreads);

namespace pp {
std::string last_error;
void set_error(const std::string &msg) { last_error = msg; }
}  // namespace pp

int main(int argc, char **argv)
{
    const std::string path = argc > 1 ? argv[1] : "/tmp/ppafm_asan.xsf";
    const int nx = 7, ny = 5, nz = 6;
    std::vector<double> g(nx * ny * nz * 2);
    unsigned s = 7;
    for (double &v : g) {
        s = s * 1664525u + 1013904223u;
        v = ((s >> 8) * (1.0 / 16777216.0) - 0.5) * 1e3;
    }
    g[3] = NAN;
    g[4] = INFINITY;
    g[5] = -0.0;
    g[6] = 5e-320;
    for (int nth : {1, 4}) {
        // contiguous x-fastest, then every second value (a strided view)
        if (ppafm_xsf_write_data(path.c_str(), 0, g.data(), nx, ny, nz, 1, nx, (long long)nx * ny, nth) != 0 ||
            ppafm_xsf_write_data(path.c_str(), 1, g.data(), nx, ny, nz, 2, 2 * nx, 2LL * nx * ny, nth) != 0) {
            std::printf("write failed: %s\n", pp::last_error.c_str());
            return 1;
        }
    }
    // the first data block (the END lines after it are not numbers)
    const long long count = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    for (int nth : {1, 3}) {
        std::vector<double> out(count);
        if (ppafm_read_numbers(path.c_str(), 0, count, out.data(), nth) != 0) {
            std::printf("read failed: %s\n", pp::last_error.c_str());
            return 1;
        }
    }
    // asking for numbers past the block is an error, not an overrun
    std::vector<double> big(count + 5);
    const int rc = ppafm_read_numbers(path.c_str(), 0, count + 5, big.data(), 2);
    std::remove(path.c_str());
    std::printf("field_io sanitizer driver: ok (over-long read rc %d: %s)\n", rc, pp::last_error.c_str());
    return 0;
}
