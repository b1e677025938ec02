// Grid text I/O of ppafm's file formats, multithreaded host code.
//
//  * XSF data blocks as ppafm/io.py:575-599 (saveXSFData) writes them: the
//    grid extended by one periodic plane along each axis (the first plane
//    repeated last), x fastest, one value per line formatted "%10.5e", then
//    the two END lines.  Values are formatted with std::to_chars
//    (scientific, 5 digits after the point), which rounds exactly like the
//    reference's Python %-formatting; NaN and infinities are padded to width
//    10 as there.
//  * Whitespace-separated numbers after a given number of header lines as
//    GridUtils.cpp:32-71 (ReadNumsUpTo_C) reads them for loadXSFData and
//    loadCUBE: lines are skipped as fgets with a 5000-byte buffer skips them,
//    numbers parsed with std::from_chars (correctly rounded, as strtod).
//
// Both split the work into contiguous ranges, one per thread, and join the
// results in order, so the bytes written and the values read do not depend
// on the thread count.

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace pp {
void set_error(const std::string &msg);
}

namespace {

int thread_count(int req, long long work, long long per)
{
    int hw = (int)std::thread::hardware_concurrency();
    if (hw <= 0) hw = 1;
    int t = req > 0 ? req : std::min(hw, 32);
    const long long cap = std::max(1LL, work / per);
    return (int)std::max(1LL, std::min((long long)t, cap));
}

template <class F>
void parallel_for(int nth, F &&f)
{
    if (nth == 1) {
        f(0);
        return;
    }
    std::vector<std::thread> ts;
    ts.reserve(nth);
    for (int t = 0; t < nth; t++) ts.emplace_back(f, t);
    for (auto &th : ts) th.join();
}

// One value as "%10.5e\n"; returns the bytes written (at most 14).
inline size_t fmt_e5(char *p, double v)
{
    if (std::isnan(v)) {
        std::memcpy(p, "       nan\n", 11);
        return 11;
    }
    if (std::isinf(v)) {
        std::memcpy(p, v > 0 ? "       inf\n" : "      -inf\n", 11);
        return 11;
    }
    // every finite value takes at least 11 characters, so width 10 never pads
    const auto r = std::to_chars(p, p + 24, v, std::chars_format::scientific, 5);
    *r.ptr = '\n';
    return (size_t)(r.ptr - p) + 1;
}

constexpr long long kWriteRound = 1LL << 23;  // values formatted per round (bounds the buffers)

}  // namespace

extern "C" {

int ppafm_xsf_write_data(const char *fname, int append, const double *data, int nx, int ny, int nz, long long sx,
                         long long sy, long long sz, int nthreads)
{
    if (!fname || !data || nx <= 0 || ny <= 0 || nz <= 0) {
        pp::set_error("ppafm_xsf_write_data: null argument or empty grid");
        return -2;
    }
    FILE *f = std::fopen(fname, append ? "ab" : "wb");
    if (!f) {
        pp::set_error(std::string("cannot open ") + fname + " for writing");
        return -1;
    }
    const long long X = nx + 1, Y = ny + 1, XY = X * Y, N = XY * (nz + 1);
    std::vector<std::vector<char>> bufs;
    bool ok = true;
    for (long long r0 = 0; r0 < N && ok; r0 += kWriteRound) {
        const long long r1 = std::min(N, r0 + kWriteRound);
        const int nth = thread_count(nthreads, r1 - r0, 1 << 16);
        bufs.assign(nth, {});
        parallel_for(nth, [&](int t) {
            const long long a = r0 + (r1 - r0) * t / nth, b = r0 + (r1 - r0) * (t + 1) / nth;
            std::vector<char> &buf = bufs[t];
            buf.resize((size_t)(b - a) * 16);
            size_t n = 0;
            for (long long k = a; k < b; k++) {
                long long iz = k / XY, rem = k - iz * XY, iy = rem / X, ix = rem - iy * X;
                if (ix == nx) ix = 0;
                if (iy == ny) iy = 0;
                if (iz == nz) iz = 0;
                n += fmt_e5(buf.data() + n, data[ix * sx + iy * sy + iz * sz]);
            }
            buf.resize(n);
        });
        for (auto &b : bufs)
            if (!b.empty() && std::fwrite(b.data(), 1, b.size(), f) != b.size()) ok = false;
    }
    if (ok && std::fputs("   END_DATAGRID_3D\nEND_BLOCK_DATAGRID_3D\n", f) < 0) ok = false;
    if (std::fclose(f) != 0) ok = false;
    if (!ok) {
        pp::set_error(std::string("write to ") + fname + " failed");
        return -1;
    }
    return 0;
}

int ppafm_read_numbers(const char *fname, int noline, long long count, double *out, int nthreads)
{
    if (!fname || (count > 0 && !out) || count < 0 || noline < 0) {
        pp::set_error("ppafm_read_numbers: bad argument");
        return -2;
    }
    FILE *f = std::fopen(fname, "rb");
    if (!f) {
        pp::set_error(std::string("cannot open ") + fname);
        return -1;
    }
    std::string text;
    {
        std::vector<char> chunk(1 << 22);
        size_t got;
        while ((got = std::fread(chunk.data(), 1, chunk.size(), f)) > 0) text.append(chunk.data(), got);
        std::fclose(f);
    }
    const char *s = text.data();
    const size_t size = text.size();
    size_t pos = 0;
    // fgets(line, 5000, f): up to 4999 bytes, through the first newline
    for (int i = 0; i < noline && pos < size; i++) {
        const size_t lim = std::min(size, pos + 4999);
        const void *nl = std::memchr(s + pos, '\n', lim - pos);
        pos = nl ? (size_t)((const char *)nl - s) + 1 : lim;
    }
    auto is_ws = [](char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; };
    const size_t body = size - pos;
    const int nth = thread_count(nthreads, (long long)body, 1 << 20);
    // chunk boundaries moved forward to whitespace so no token is split
    std::vector<size_t> cut(nth + 1);
    cut[0] = pos;
    cut[nth] = size;
    for (int t = 1; t < nth; t++) {
        size_t c = pos + body * t / nth;
        c = std::max(c, cut[t - 1]);
        while (c < size && !is_ws(s[c])) c++;
        cut[t] = c;
    }
    std::vector<long long> ntok(nth + 1, 0);
    parallel_for(nth, [&](int t) {
        long long n = 0;
        bool in = false;
        for (size_t i = cut[t]; i < cut[t + 1]; i++) {
            const bool w = is_ws(s[i]);
            if (!w && !in) n++;
            in = !w;
        }
        ntok[t + 1] = n;
    });
    for (int t = 0; t < nth; t++) ntok[t + 1] += ntok[t];
    if (ntok[nth] < count) {
        pp::set_error(std::string(fname) + ": expected " + std::to_string(count) + " numbers after " +
                      std::to_string(noline) + " lines, found " + std::to_string(ntok[nth]) + " fields");
        return -1;
    }
    std::vector<long long> bad(nth, -1);
    parallel_for(nth, [&](int t) {
        long long k = ntok[t];
        size_t i = cut[t];
        const size_t e = cut[t + 1];
        while (i < e && k < count) {
            while (i < e && is_ws(s[i])) i++;
            if (i >= e) break;
            size_t j = i;
            while (j < e && !is_ws(s[j])) j++;
            const char *a = s + i;
            if (*a == '+') a++;
            double v = 0;
            const auto r = std::from_chars(a, s + j, v, std::chars_format::general);
            if (r.ec != std::errc() || r.ptr != s + j) {
                if (r.ec == std::errc::result_out_of_range) {
                    // strtod gives +-HUGE_VAL or a subnormal/zero here
                    v = std::strtod(std::string(a, s + j).c_str(), nullptr);
                } else {
                    bad[t] = k;
                    return;
                }
            }
            out[k++] = v;
            i = j;
        }
    });
    for (int t = 0; t < nth; t++)
        if (bad[t] >= 0) {
            pp::set_error(std::string(fname) + ": field " + std::to_string(bad[t]) + " after the header is not a number");
            return -1;
        }
    return 0;
}

}  // extern "C"
