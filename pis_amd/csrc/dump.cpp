// Trajectory text writer: DumpTraj::write_step (src/writers/dump_traj.rs:33-74) as host code.
//
// A block is the LAMMPS text frame the reference writes: timestep, atom count, the three
// diagonal box bounds "0 h_kk", then one "id type x y z" line per atom in original order.
// Every f64 is printed as Rust's `{}` prints it: the shortest digit string that reads back
// to the same value, expanded positionally (never an exponent), "-0" for negative zero,
// "inf"/"-inf" and "NaN".  The shortest digits come from std::to_chars (C++17, shortest
// round trip); only the placement of the decimal point is done here.
//
// The atom lines are formatted by up to `threads` worker threads into private buffers,
// which are then written to the file descriptor in atom order, so the bytes are the same
// whatever the thread count.
#include <algorithm>
#include <charconv>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "engine.h"

namespace {

// Appends Rust's Display text of x.  At most 330 characters (5e-324 has 324 zeros after
// the point before its digit; 1.8e308 has 309 integer digits).
char* put_f64(char* p, double x)
{
    if (std::isnan(x)) {
        std::memcpy(p, "NaN", 3);
        return p + 3;
    }
    if (std::isinf(x)) {
        if (x < 0) *p++ = '-';
        std::memcpy(p, "inf", 3);
        return p + 3;
    }
    char sci[40];
    const auto r = std::to_chars(sci, sci + sizeof sci, x, std::chars_format::scientific);
    const char* s = sci;
    if (*s == '-') {
        *p++ = '-';
        ++s;
    }
    // s = d[.ddd]e(+|-)XX
    char dig[24];
    int nd = 0;
    while (*s != 'e') {
        if (*s != '.') dig[nd++] = *s;
        ++s;
    }
    ++s;
    const bool neg_exp = (*s == '-');
    ++s;
    int e = 0;
    while (s < r.ptr) e = e * 10 + (*s++ - '0');
    if (neg_exp) e = -e;
    const int point = e + 1;  // digits before the decimal point
    if (point <= 0) {
        *p++ = '0';
        *p++ = '.';
        std::memset(p, '0', (size_t)-point);
        p += -point;
        std::memcpy(p, dig, (size_t)nd);
        p += nd;
    } else if (point >= nd) {
        std::memcpy(p, dig, (size_t)nd);
        p += nd;
        std::memset(p, '0', (size_t)(point - nd));
        p += point - nd;
    } else {
        std::memcpy(p, dig, (size_t)point);
        p += point;
        *p++ = '.';
        std::memcpy(p, dig + point, (size_t)(nd - point));
        p += nd - point;
    }
    return p;
}

char* put_u64(char* p, uint64_t v)
{
    return std::to_chars(p, p + 24, v).ptr;
}

constexpr size_t kLineMax = 24 + 1 + 12 + 3 * (1 + 330) + 1;

// "id type x y z\n" for atoms [a, b) into out (grown as needed); false if memory ran out
// (a worker thread must not let an exception escape).
bool format_atoms_into(const double* pos, const uint32_t* type_ids, int64_t a, int64_t b, std::string& out);

void format_atoms(const double* pos, const uint32_t* type_ids, int64_t a, int64_t b, std::string& out)
{
    out.clear();
    size_t used = 0;
    out.resize((size_t)(b - a) * 64 + kLineMax);
    for (int64_t i = a; i < b; ++i) {
        if (out.size() - used < kLineMax) out.resize(out.size() * 2 + kLineMax);
        char* p = &out[used];
        char* q = p;
        q = put_u64(q, (uint64_t)i + 1);
        *q++ = ' ';
        q = put_u64(q, type_ids[i]);
        for (int c = 0; c < 3; ++c) {
            *q++ = ' ';
            q = put_f64(q, pos[3 * i + c]);
        }
        *q++ = '\n';
        used += (size_t)(q - p);
    }
    out.resize(used);
}

bool format_atoms_into(const double* pos, const uint32_t* type_ids, int64_t a, int64_t b, std::string& out)
{
    try {
        format_atoms(pos, type_ids, a, b, out);
        return true;
    } catch (...) {
        out.clear();
        return false;
    }
}

bool write_all(int fd, const char* p, size_t len)
{
    while (len > 0) {
        const ssize_t w = ::write(fd, p, len);
        if (w < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        p += w;
        len -= (size_t)w;
    }
    return true;
}

}  // namespace

extern "C" pis_status pis_format_f64(double x, char* out, size_t cap, size_t* len)
{
    char tmp[400];
    char* e = put_f64(tmp, x);
    const size_t n = (size_t)(e - tmp);
    if (len) *len = n;
    if (!out || cap < n + 1) return pis::fail(PIS_ERR_INVALID_ARGUMENT, "output buffer too small");
    std::memcpy(out, tmp, n);
    out[n] = '\0';
    return PIS_OK;
}

extern "C" pis_status pis_dump_write_block(int fd, uint64_t step, uint64_t n_atoms, const double* pos,
                                           const uint32_t* type_ids, const double* h, int32_t threads)
{
    if (fd < 0 || !h || (n_atoms > 0 && (!pos || !type_ids)))
        return pis::fail(PIS_ERR_INVALID_ARGUMENT, "pis_dump_write_block: bad argument");
    std::string head;
    head.reserve(256 + 3 * 340);
    char buf[400];
    head += "ITEM: TIMESTEP\n";
    head.append(buf, put_u64(buf, step));
    head += "\nITEM: NUMBER OF ATOMS\n";
    head.append(buf, put_u64(buf, n_atoms));
    head += "\nITEM: BOX BOUNDS pp pp pp\n";
    for (int k = 0; k < 3; ++k) {
        head += "0 ";
        head.append(buf, put_f64(buf, h[4 * k]));  // h(k,k), column-major
        head += '\n';
    }
    head += "ITEM: ATOMS id type x y z\n";

    const int64_t n = (int64_t)n_atoms;
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    nt = std::max(1, std::min(nt, 64));
    const int64_t per = std::max<int64_t>(4096, (n + nt - 1) / std::max(1, nt));
    const int chunks = (int)std::max<int64_t>(1, (n + per - 1) / per);
    std::vector<std::string> parts((size_t)chunks);
    std::vector<char> ok((size_t)chunks, 0);
    std::vector<std::thread> pool;
    bool started = true;
    try {
        pool.reserve((size_t)chunks - 1);
        for (int c = 1; c < chunks; ++c)
            pool.emplace_back([&, c] {
                ok[c] = format_atoms_into(pos, type_ids, c * per, std::min(n, (c + 1) * per), parts[c]);
            });
    } catch (...) {  // a thread did not start: the started ones are joined below
        started = false;
    }
    ok[0] = format_atoms_into(pos, type_ids, 0, std::min(n, per), parts[0]);
    for (auto& t : pool) t.join();
    if (!started || std::find(ok.begin(), ok.end(), 0) != ok.end())
        return pis::fail(PIS_ERR_IO, "trajectory formatting failed (threads or memory)");
    if (!write_all(fd, head.data(), head.size()))
        return pis::fail(PIS_ERR_IO, std::string("Failed to write trajectory block: ") + std::strerror(errno));
    for (const auto& s : parts)
        if (!write_all(fd, s.data(), s.size()))
            return pis::fail(PIS_ERR_IO, std::string("Failed to write trajectory block: ") + std::strerror(errno));
    return PIS_OK;
}
