// nsp_obj.cpp -- reading and writing OBJ / OBJN files (host code).
//
// Reading follows FileObj<N> (src/obj/file/file_load.cpp:946-1144, 1402-1430)
// and obj_file_dimension_and_type (src/obj/file/obj_file.cpp:144-199):
//   * the dimension comes from the extension: "obj" is 3, "objN" is N (N >= 3);
//   * '#' starts a comment; the first word of a line is its keyword;
//   * "v": N floats; "vt": N-1 floats, or N with the last one 0; "vn": N
//     floats, normalised in float; "f": vertex groups "v", "v/t", "v//n",
//     "v/t/n"; other keywords are skipped.  Floats are read as strtof reads
//     them, and a float out of range or not finite ends the list;
//   * a face line with G groups is fanned out into G - (N - 1) facets (group 0
//     with groups i+1 .. i+N-1), at most 5 facets for N = 3 and 1 otherwise;
//     texture and normal indices are given on all groups or none; index 0 is
//     an error; positive indices are 1-based, negative ones count back from
//     the number of items read before the line;
//   * then every index is range-checked against the file's totals, the
//     facets' bounding box must have positive extent along every axis, and
//     for N = 3 facets whose edge vectors are parallel are removed.
// The file is split into line chunks parsed on several threads; per-chunk item
// counts give, by a prefix sum, the running counts each chunk needs to resolve
// negative indices, so both passes run in parallel.
//
// Writing follows save_obj_geometry_to_file (src/obj/file/file_save.cpp:
// 36-306): an optional "# " comment, vertices scaled into [-1, 1] about the
// centre of the bounding box of the facets' vertices (proportions kept),
// "%12.9f" per coordinate, normals normalised in double, facets "f a b c" or
// "f a//n b//n c//n" (1-based); for N = 3 a facet whose vertex normals all
// point against its edge cross product is written with vertices 1 and 2
// swapped.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "nsp.h"

namespace {

thread_local std::string g_obj_error;

struct ObjError {
    std::string msg;
};

[[noreturn]] void fail(const std::string& m) { throw ObjError{m}; }

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

// file_extension (src/com/file/file_sys.cpp:61): text after the last '.' of
// the base name
std::string extension_of(const std::string& path)
{
    const size_t slash = path.find_last_of('/');
    const std::string base = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = base.find_last_of('.');
    return dot == std::string::npos ? std::string() : base.substr(dot + 1);
}

int dimension_of(const std::string& path)
{
    const std::string e = extension_of(path);
    if (e.empty()) fail("No OBJ file extension found");
    if (e == "obj") return 3;
    if (e.rfind("obj", 0) != 0) fail("Unsupported OBJ file format " + e);
    const std::string s = e.substr(3);
    if (s.empty() || !std::all_of(s.begin(), s.end(), is_digit)) fail("Wrong dimension number string \"" + s + "\"");
    if (s.size() > 6) fail("Failed to read dimension number from string \"" + s + "\"");
    const int d = std::atoi(s.c_str());
    if (d < 3) fail("Wrong dimension number " + std::to_string(d));
    return d;
}

// Exact decimal-to-float conversion for the common case.  A token of at most
// 19 significant digits m and decimal exponent |e| <= 22 with m < 2^53 gives
// d = m * 10^e (or m / 10^-e) rounded once to double; rounding d to float then
// gives the correctly rounded float of the decimal value unless d lies exactly
// on a midpoint between two floats (no other float midpoint can lie between
// the decimal value and d, since midpoints are doubles).  Returns false, and
// leaves the decision to strtof, for anything else: midpoints, results that
// are zero from non-zero digits, subnormal or infinite, more digits, hex,
// "inf", "nan" or no digits at all.  Accepted tokens are consumed exactly as
// strtof consumes them (leading white space, sign, digits, '.', exponent).
bool fast_float(const char* p, const char* e, float* out, const char** end)
{
    static constexpr double P10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                       1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const char* q = p;
    while (q < e && is_space(*q)) ++q;
    bool neg = false;
    if (q < e && (*q == '+' || *q == '-')) {
        neg = *q == '-';
        ++q;
    }
    if (q + 1 < e && q[0] == '0' && (q[1] == 'x' || q[1] == 'X')) return false;
    uint64_t m = 0;
    int digits = 0, e10 = 0;
    bool any = false, dropped = false;
    for (; q < e && is_digit(*q); ++q) {
        any = true;
        if (digits < 19) {
            m = m * 10 + (uint64_t)(*q - '0');
            digits += m != 0;
        } else {
            ++e10;
            dropped |= *q != '0';
        }
    }
    if (q < e && *q == '.') {
        for (++q; q < e && is_digit(*q); ++q) {
            any = true;
            if (digits < 19) {
                m = m * 10 + (uint64_t)(*q - '0');
                digits += m != 0;
                --e10;
            } else {
                dropped |= *q != '0';
            }
        }
    }
    if (!any || dropped) return false;
    if (q < e && (*q == 'e' || *q == 'E')) {
        const char* r = q + 1;
        bool eneg = false;
        if (r < e && (*r == '+' || *r == '-')) {
            eneg = *r == '-';
            ++r;
        }
        if (r < e && is_digit(*r)) {
            int x = 0;
            for (; r < e && is_digit(*r); ++r)
                if (x < 100000) x = x * 10 + (*r - '0');
            e10 += eneg ? -x : x;
            q = r;
        }
    }
    if (m == 0) {
        *out = neg ? -0.0f : 0.0f;
        *end = q;
        return true;
    }
    if (m >= ((uint64_t)1 << 53) || e10 < -22 || e10 > 22) return false;
    const double d = e10 >= 0 ? (double)m * P10[e10] : (double)m / P10[-e10];
    const float f = (float)d;
    if (!(f >= FLT_MIN) || f > FLT_MAX) return false;
    uint64_t bits;
    std::memcpy(&bits, &d, sizeof bits);
    if ((bits & (((uint64_t)1 << 29) - 1)) == ((uint64_t)1 << 28)) return false;
    *out = neg ? -f : f;
    *end = q;
    return true;
}

// string_to_floats (src/obj/file/file_load.cpp:448-498): values read as strtof
// reads them, in sequence from p up to the end e of the line's text, stopping
// at the first value that is missing, out of range or not finite
int read_floats(const char* p, const char* e, float* out, int count, std::string& tmp)
{
    int k = 0;
    while (k < count) {
        const char* end;
        float v;
        if (!fast_float(p, e, &v, &end)) {
            tmp.assign(p, e);  // strtof on a terminated copy of the rest of the line
            char* te;
            errno = 0;
            v = std::strtof(tmp.c_str(), &te);
            if (te == tmp.c_str() || errno == ERANGE || !std::isfinite(v)) break;
            end = p + (te - tmp.c_str());
        }
        out[k++] = v;
        p = end;
    }
    return k;
}

// read_integer (src/obj/file/file_load.cpp:272-296): optional '-', digits
bool read_integer(const char* e, const char** pos, int32_t* value)
{
    const char* b = *pos;
    const bool neg = b < e && *b == '-';
    if (neg) ++b;
    const char* q = b;
    int64_t sum = 0;
    while (q < e && is_digit(*q)) {
        sum = sum * 10 + (*q - '0');
        if (sum > INT32_MAX) fail("Facet index too large");
        ++q;
    }
    if (q == b) return false;
    *value = (int32_t)(neg ? -sum : sum);
    *pos = q;
    return true;
}

// normalize (src/com/vec.h:259-290): sum of squares with fma, sqrt, divide
template <typename T>
void normalize_fma(T* v, int n)
{
    T r = v[0] * v[0];
    for (int i = 1; i < n; ++i) r = std::fma(v[i], v[i], r);
    const T len = std::sqrt(r);
    for (int i = 0; i < n; ++i) v[i] = v[i] / len;
}

void run_threads(int T, const std::function<void(int)>& fn)
{
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(fn, t);
    fn(0);
    for (auto& th : pool) th.join();
}

// the file's bytes, mapped read-only
struct Mapped {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    explicit Mapped(const char* path)
    {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) fail(std::string("Failed to open file ") + path);
        struct stat st;
        if (::fstat(fd, &st) != 0) fail(std::string("Failed to read file ") + path);
        n = (size_t)st.st_size;
        if (n == 0) return;
        void* m = ::mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) fail(std::string("Failed to read file ") + path);
        p = (const char*)m;
    }
    ~Mapped()
    {
        if (p) ::munmap((void*)p, n);
        if (fd >= 0) ::close(fd);
    }
};

struct Chunk {
    const char* begin = nullptr;
    const char* end = nullptr;
    std::vector<float> v, vn;
    int64_t nv = 0, nvt = 0, nvn = 0;  // items of this chunk
    // per facet: N (vertex, texture, normal) indices as read, and the chunk's
    // running (v, vt, vn) counts at its line
    std::vector<std::array<int32_t, 3>> g;
    std::vector<std::array<int32_t, 3>> at;
    int64_t lines = 0;
    std::string error;
    // after index correction: the box of the chunk's facet vertices, and how
    // many of its triangles have parallel edge vectors (n = 3)
    std::vector<float> lo, hi;
    int64_t flat = 0;
};

void parse_chunk(Chunk& c, int N)
{
    const int max_facets = N == 3 ? 5 : 1;
    const int max_groups = max_facets + N - 1;
    {
        // room for a line of n values or one facet per line
        int64_t lines = 0;
        for (const char* p = c.begin; p < c.end; ++lines) {
            const char* le = (const char*)std::memchr(p, '\n', (size_t)(c.end - p));
            p = le ? le + 1 : c.end;
        }
        c.g.reserve((size_t)(lines * N));
        c.at.reserve((size_t)lines);
    }
    std::vector<std::array<int32_t, 3>> grp((size_t)max_groups);
    std::vector<float> tmp((size_t)N + 1);
    std::string text;
    const char* p = c.begin;
    while (p < c.end) {
        const char* le = (const char*)std::memchr(p, '\n', (size_t)(c.end - p));
        if (!le) le = c.end;
        const char* ce = (const char*)std::memchr(p, '#', (size_t)(le - p));
        if (!ce) ce = le;
        const char* q = p;
        while (q < ce && is_space(*q)) ++q;
        const char* kw = q;
        while (q < ce && !is_space(*q)) ++q;
        const size_t kl = (size_t)(q - kw);
        const bool v = kl == 1 && kw[0] == 'v';
        const bool vt = kl == 2 && kw[0] == 'v' && kw[1] == 't';
        const bool vn = kl == 2 && kw[0] == 'v' && kw[1] == 'n';
        const bool f = kl == 1 && kw[0] == 'f';
        if (v || vn) {
            if (read_floats(q, ce, tmp.data(), N, text) != N)
                fail("Error read " + std::to_string(N) + " floating points of float type");
            if (v) {
                c.v.insert(c.v.end(), tmp.begin(), tmp.begin() + N);
                ++c.nv;
            } else {
                normalize_fma(tmp.data(), N);
                c.vn.insert(c.vn.end(), tmp.begin(), tmp.begin() + N);
                ++c.nvn;
            }
        } else if (vt) {
            const int k = read_floats(q, ce, tmp.data(), N, text);
            if (k != N - 1 && k != N)
                fail("Error read " + std::to_string(N - 1) + " or " + std::to_string(N) +
                     " floating points of float type");
            if (k == N && tmp[(size_t)N - 1] != 0)
                fail(std::to_string(N) + "-dimensional textures are not supported");
            ++c.nvt;
        } else if (f) {
            // read_digit_groups, read_facets (src/obj/file/file_load.cpp:300-446)
            int count = 0;
            const char* r = q;
            while (true) {
                while (r < ce && is_space(*r)) ++r;
                if (r >= ce) break;
                if (count >= max_groups)
                    fail("Found too many facet vertices " + std::to_string(count + 1) + " (max supported = " +
                         std::to_string(max_groups) + ")");
                std::array<int32_t, 3>& idx = grp[(size_t)count];
                if (!read_integer(ce, &r, &idx[0])) fail("Error read facet vertex first number");
                if (idx[0] == 0) fail("Zero facet index");
                for (int a = 1; a < 3; ++a) {
                    idx[a] = 0;
                    if (r >= ce || is_space(*r)) continue;
                    if (*r != '/') fail("Error read facet vertex number");
                    ++r;
                    if (r >= ce || is_space(*r)) continue;
                    if (read_integer(ce, &r, &idx[a])) {
                        if (idx[a] == 0) fail("Zero facet index");
                    } else {
                        idx[a] = 0;
                    }
                }
                ++count;
            }
            if (count < N)
                fail("Error facet vertex count " + std::to_string(count) + " (min = " + std::to_string(N) + ")");
            int tex = 0, nrm = 0;
            for (int i = 0; i < count; ++i) {
                tex += grp[(size_t)i][1] != 0;
                nrm += grp[(size_t)i][2] != 0;
            }
            if (!(tex == 0 || tex == count)) fail("Inconsistent facet texture indices");
            if (!(nrm == 0 || nrm == count)) fail("Inconsistent facet normal indices");
            if (c.nv > INT32_MAX || c.nvt > INT32_MAX || c.nvn > INT32_MAX) fail("OBJ file too large");
            for (int fi = 0; fi < count - (N - 1); ++fi) {
                for (int k = 0; k < N; ++k) c.g.push_back(grp[(size_t)(k == 0 ? 0 : fi + k)]);
                c.at.push_back({(int32_t)c.nv, (int32_t)c.nvt, (int32_t)c.nvn});
            }
        }
        ++c.lines;
        p = le + 1;
    }
}

// center_and_length_for_min_max (src/obj/alg/alg.h:36-70) on the box of the facets' vertices
void check_box(const float* mn, const float* mx, int N)
{
    long double len2 = 0;
    for (int i = 0; i < N; ++i) {
        if (mn[i] >= mx[i]) fail("Object size error");
        const float d = mx[i] - mn[i];
        if (!std::isfinite(mn[i] + d / 2)) fail("Object center is not finite");
        len2 += (long double)d * d;
    }
    if (!std::isfinite((double)std::sqrt(len2))) fail("Object length is not finite");
}

// facet_dimension_is_correct (src/obj/file/file_load.cpp:688-716)
bool triangle_has_area(const float* a, const float* b, const float* c)
{
    double e0[3], e1[3];
    for (int i = 0; i < 3; ++i) {
        e0[i] = (double)(b[i] - a[i]);
        e1[i] = (double)(c[i] - a[i]);
    }
    return e0[1] * e1[2] - e0[2] * e1[1] != 0 || e0[0] * e1[2] - e0[2] * e1[0] != 0 ||
           e0[0] * e1[1] - e0[1] * e1[0] != 0;
}

}  // namespace

struct nsp_obj {
    int n = 0;
    int64_t V = 0, VN = 0, VT = 0, F = 0;
    std::unique_ptr<float[]> vertices, normals;
    std::unique_ptr<int32_t[]> facets, facet_normals;
};

extern "C" {

const char* nsp_obj_last_error(void) { return g_obj_error.c_str(); }

int nsp_obj_read(const char* path, int threads, nsp_obj** out)
{
    if (!path || !out) {
        g_obj_error = "invalid argument";
        return NSP_E_INVALID;
    }
    *out = nullptr;
    try {
        const int N = dimension_of(path);
        const Mapped file(path);
        // chunks at line boundaries; automatic thread counts give each chunk at
        // least 64 KiB
        const size_t size = file.n;
        int T = threads > 0 ? std::min(threads, 64)
                            : std::min({(int)std::max(1u, std::thread::hardware_concurrency()), 64,
                                        1 + (int)std::min<size_t>(size >> 16, 64)});
        std::vector<Chunk> chunks((size_t)T);
        const char* base = file.p;
        const char* end = base + size;
        const char* cur = base;
        for (int t = 0; t < T; ++t) {
            const char* e = t == T - 1 ? end : base + size * (size_t)(t + 1) / (size_t)T;
            if (e < cur) e = cur;
            while (e < end && *e != '\n') ++e;
            if (e < end) ++e;
            chunks[(size_t)t].begin = cur;
            chunks[(size_t)t].end = e;
            cur = e;
        }
        run_threads(T, [&](int t) {
            try {
                parse_chunk(chunks[(size_t)t], N);
            } catch (const ObjError& e) {
                chunks[(size_t)t].error = e.msg;
            } catch (const std::bad_alloc&) {
                chunks[(size_t)t].error = "host allocation failed";
            }
        });
        // the first error in file order (the chunks before it were read whole);
        // lines are numbered from 0
        int64_t line0 = 0;
        for (const Chunk& c : chunks) {
            if (!c.error.empty()) fail("Line " + std::to_string(line0 + c.lines) + ": " + c.error);
            line0 += c.lines;
        }
        auto obj = std::make_unique<nsp_obj>();
        obj->n = N;
        std::vector<std::array<int64_t, 4>> start((size_t)T);  // v, vt, vn, facets before each chunk
        std::array<int64_t, 4> tot = {0, 0, 0, 0};
        for (int t = 0; t < T; ++t) {
            const Chunk& c = chunks[(size_t)t];
            start[(size_t)t] = tot;
            tot[0] += c.nv;
            tot[1] += c.nvt;
            tot[2] += c.nvn;
            tot[3] += (int64_t)c.at.size();
        }
        const int64_t V = tot[0], VT = tot[1], VN = tot[2], F = tot[3];
        if (F == 0) fail("No facets found in OBJ file");
        if (V > INT32_MAX || VN > INT32_MAX || F > INT32_MAX) fail("OBJ file too large");
        obj->V = V;
        obj->VT = VT;
        obj->VN = VN;
        obj->F = F;
        obj->vertices.reset(new float[(size_t)(V * N)]);
        obj->normals.reset(new float[(size_t)(VN * N)]);
        obj->facets.reset(new int32_t[(size_t)(F * N)]);
        obj->facet_normals.reset(new int32_t[(size_t)(F * N)]);
        run_threads(T, [&](int t) {
            const Chunk& c = chunks[(size_t)t];
            std::copy(c.v.begin(), c.v.end(), obj->vertices.get() + start[(size_t)t][0] * N);
            std::copy(c.vn.begin(), c.vn.end(), obj->normals.get() + start[(size_t)t][2] * N);
        });
        // correct_indices, then check_facet_indices (src/obj/file/file_load.cpp:1046-1065,
        // 851-897); the box of each chunk's facet vertices and its flat triangles
        run_threads(T, [&](int t) {
            Chunk& c = chunks[(size_t)t];
            const std::array<int64_t, 4>& s0 = start[(size_t)t];
            c.lo.assign((size_t)N, std::numeric_limits<float>::max());
            c.hi.assign((size_t)N, std::numeric_limits<float>::lowest());
            const float* vtx = obj->vertices.get();
            for (size_t k = 0; k < c.at.size(); ++k) {
                const int64_t nv = s0[0] + c.at[k][0], nt = s0[1] + c.at[k][1], nn = s0[2] + c.at[k][2];
                const int64_t fo = (s0[3] + (int64_t)k) * N;
                for (int i = 0; i < N; ++i) {
                    const std::array<int32_t, 3>& g = c.g[k * N + (size_t)i];
                    const int64_t a = g[0], b = g[1], d = g[2];
                    const int64_t va = a > 0 ? a - 1 : nv + a;
                    const int64_t tb = b > 0 ? b - 1 : (b < 0 ? nt + b : -1);
                    const int64_t nd = d > 0 ? d - 1 : (d < 0 ? nn + d : -1);
                    if (va < 0 || va >= V) {
                        c.error = "Vertex index " + std::to_string(va) + " is out of bounds [0, " +
                                  std::to_string(V) + ")";
                        return;
                    }
                    if (b != 0 && (tb < 0 || tb >= VT)) {
                        c.error = "Texture coordinate index " + std::to_string(tb) + " is out of bounds [0, " +
                                  std::to_string(VT) + ")";
                        return;
                    }
                    if (d != 0 && (nd < 0 || nd >= VN)) {
                        c.error = "Normal index " + std::to_string(nd) + " is out of bounds [0, " +
                                  std::to_string(VN) + ")";
                        return;
                    }
                    obj->facets[(size_t)(fo + i)] = (int32_t)va;
                    obj->facet_normals[(size_t)(fo + i)] = (int32_t)nd;
                    for (int j = 0; j < N; ++j) {
                        const float x = vtx[va * N + j];
                        c.lo[(size_t)j] = std::min(c.lo[(size_t)j], x);
                        c.hi[(size_t)j] = std::max(c.hi[(size_t)j], x);
                    }
                }
                if (N == 3) {
                    const int32_t* fv = &obj->facets[(size_t)fo];
                    c.flat += !triangle_has_area(vtx + 3 * (int64_t)fv[0], vtx + 3 * (int64_t)fv[1],
                                                 vtx + 3 * (int64_t)fv[2]);
                }
            }
        });
        std::vector<float> lo((size_t)N, std::numeric_limits<float>::max()),
            hi((size_t)N, std::numeric_limits<float>::lowest());
        int64_t flat = 0;
        for (const Chunk& c : chunks) {
            if (!c.error.empty()) fail(c.error);
            for (int j = 0; j < N; ++j) {
                lo[(size_t)j] = std::min(lo[(size_t)j], c.lo[(size_t)j]);
                hi[(size_t)j] = std::max(hi[(size_t)j], c.hi[(size_t)j]);
            }
            flat += c.flat;
        }
        check_box(lo.data(), hi.data(), N);
        if (flat > 0) {
            // remove_facets_with_incorrect_dimension (src/obj/file/file_load.cpp:902-940), n = 3
            if (flat == F) fail("No 2-facets found in OBJ-3 file");
            int32_t* fac = obj->facets.get();
            int32_t* fnr = obj->facet_normals.get();
            const float* vv = obj->vertices.get();
            int64_t kept = 0;
            std::fill(lo.begin(), lo.end(), std::numeric_limits<float>::max());
            std::fill(hi.begin(), hi.end(), std::numeric_limits<float>::lowest());
            for (int64_t k = 0; k < F; ++k) {
                const int32_t* fv = fac + 3 * k;
                if (!triangle_has_area(vv + 3 * (int64_t)fv[0], vv + 3 * (int64_t)fv[1], vv + 3 * (int64_t)fv[2]))
                    continue;
                for (int i = 0; i < 3; ++i) {
                    const int32_t x = fac[3 * k + i];
                    fac[3 * kept + i] = x;
                    fnr[3 * kept + i] = fnr[3 * k + i];
                    for (int j = 0; j < 3; ++j) {
                        lo[(size_t)j] = std::min(lo[(size_t)j], vv[3 * (int64_t)x + j]);
                        hi[(size_t)j] = std::max(hi[(size_t)j], vv[3 * (int64_t)x + j]);
                    }
                }
                ++kept;
            }
            obj->F = kept;
            check_box(lo.data(), hi.data(), N);
        }
        *out = obj.release();
        return NSP_OK;
    } catch (const ObjError& e) {
        g_obj_error = e.msg;
        return NSP_E_INVALID;
    } catch (const std::bad_alloc&) {
        g_obj_error = "host allocation failed";
        return NSP_E_NOMEM;
    }
}

void nsp_obj_destroy(nsp_obj* obj) { delete obj; }

int nsp_obj_info(const nsp_obj* obj, int* n, int64_t* vertices, int64_t* facets, int64_t* normals,
                 int64_t* texcoords)
{
    if (!obj) return NSP_E_INVALID;
    if (n) *n = obj->n;
    if (vertices) *vertices = obj->V;
    if (facets) *facets = obj->F;
    if (normals) *normals = obj->VN;
    if (texcoords) *texcoords = obj->VT;
    return NSP_OK;
}

int nsp_obj_copy(const nsp_obj* obj, float* vertices, int32_t* facets, float* normals, int32_t* facet_normals)
{
    if (!obj) return NSP_E_INVALID;
    const int n = obj->n;
    if (vertices) std::copy(obj->vertices.get(), obj->vertices.get() + obj->V * n, vertices);
    if (facets) std::copy(obj->facets.get(), obj->facets.get() + obj->F * n, facets);
    if (normals) std::copy(obj->normals.get(), obj->normals.get() + obj->VN * n, normals);
    if (facet_normals) std::copy(obj->facet_normals.get(), obj->facet_normals.get() + obj->F * n, facet_normals);
    return NSP_OK;
}

int nsp_mesh_create_obj(nsp_context* ctx, const nsp_obj* obj, const double* vertex_matrix, nsp_mesh** out)
{
    if (!ctx || !obj || !out) return NSP_E_INVALID;
    const bool normals = obj->VN > 0;
    return nsp_mesh_create(ctx, obj->n, obj->vertices.get(), obj->V, obj->facets.get(), obj->F,
                           normals ? obj->normals.get() : nullptr, obj->VN,
                           normals ? obj->facet_normals.get() : nullptr, vertex_matrix, out);
}

int nsp_obj_write(const char* path, int n, const float* vertices, int64_t V, const int32_t* facets, int64_t F,
                  const float* normals, int64_t VN, const int32_t* facet_normals, const char* comment)
{
    if (!path || n < 3 || n > 64 || !vertices || !facets || V < 1 || F < 1) {
        g_obj_error = "invalid argument";
        return NSP_E_INVALID;
    }
    FILE* file = nullptr;
    try {
        // obj_file_extension_is_correct (src/obj/file/obj_file.cpp:198-201)
        const std::string e = extension_of(path);
        if (e.empty()) fail("No OBJ file extension found");
        if (e != (n == 3 ? "obj" : "obj" + std::to_string(n)) && e != "obj" + std::to_string(n))
            fail("Wrong OBJ-" + std::to_string(n) + " file name extension: " + e);
        const bool with_normals = normals && facet_normals && VN > 0;
        // a facet has normals on all of its vertices or on none
        std::vector<uint8_t> has_normal((size_t)F, 0);
        for (int64_t f = 0; f < F; ++f) {
            int k = 0;
            for (int i = 0; i < n; ++i) {
                const int32_t v = facets[f * n + i];
                if (v < 0 || v >= V) fail("Facet vertex index out of bounds");
                if (with_normals) {
                    const int32_t q = facet_normals[f * n + i];
                    if (q < -1 || q >= VN) fail("Facet normal index out of bounds");
                    k += q >= 0;
                }
            }
            if (k != 0 && k != n) fail("Inconsistent facet normal indices");
            has_normal[(size_t)f] = k == n;
        }
        // write_vertices (src/obj/file/file_save.cpp:135-172): [-1, 1] about the
        // centre of the box of the facets' vertices
        std::vector<uint8_t> used((size_t)V, 0);
        for (int64_t k = 0; k < F * n; ++k) used[(size_t)facets[k]] = 1;
        int64_t unique = 0;
        std::vector<float> mn((size_t)n, std::numeric_limits<float>::max()),
            mx((size_t)n, std::numeric_limits<float>::lowest());
        for (int64_t v = 0; v < V; ++v) {
            if (!used[(size_t)v]) continue;
            ++unique;
            for (int i = 0; i < n; ++i) {
                mn[(size_t)i] = std::min(mn[(size_t)i], vertices[v * n + i]);
                mx[(size_t)i] = std::max(mx[(size_t)i], vertices[v * n + i]);
            }
        }
        if (unique < n)
            fail("Facet unique indices count " + std::to_string(unique) + " is less than " + std::to_string(n));
        std::vector<float> delta((size_t)n), center((size_t)n);
        float max_delta = std::numeric_limits<float>::lowest();
        for (int i = 0; i < n; ++i) {
            delta[(size_t)i] = mx[(size_t)i] - mn[(size_t)i];
            max_delta = std::max(max_delta, delta[(size_t)i]);
        }
        if (max_delta == 0) fail("Object vertices are equal to each other");
        const float scale = 2 / max_delta;
        for (int i = 0; i < n; ++i) center[(size_t)i] = mn[(size_t)i] + 0.5f * delta[(size_t)i];
        // write_normals (src/obj/file/file_save.cpp:174-192): normalised in double
        std::vector<double> unit(with_normals ? (size_t)(VN * n) : 0);
        for (int64_t k = 0; with_normals && k < VN; ++k) {
            double* u = &unit[(size_t)(k * n)];
            for (int i = 0; i < n; ++i) u[i] = normals[k * n + i];
            double r = u[0] * u[0];
            for (int i = 1; i < n; ++i) r = std::fma(u[i], u[i], r);
            const double len = std::sqrt(r);
            if (len == 0) fail("Object zero length normal");
            for (int i = 0; i < n; ++i) u[i] /= len;
        }
        file = std::fopen(path, "w");
        if (!file) fail(std::string("Failed to open file ") + path);
        if (comment && *comment) {
            std::string s = "# ";
            for (const char* c = comment; *c; ++c) {
                if (*c != '\n')
                    s += *c;
                else
                    s += "\n# ";
            }
            s += '\n';
            std::fputs(s.c_str(), file);
        }
        for (int64_t v = 0; v < V; ++v) {
            std::fputs("v", file);
            for (int i = 0; i < n; ++i) {
                const float x = (vertices[v * n + i] - center[(size_t)i]) * scale;
                std::fprintf(file, " %12.9f", (double)x);
            }
            std::fputs("\n", file);
        }
        for (int64_t k = 0; with_normals && k < VN; ++k) {
            std::fputs("vn", file);
            for (int i = 0; i < n; ++i) std::fprintf(file, " %12.9f", (double)(float)unit[(size_t)(k * n + i)]);
            std::fputs("\n", file);
        }
        std::vector<int32_t> fv((size_t)n), fn((size_t)n);
        for (int64_t f = 0; f < F; ++f) {
            for (int i = 0; i < n; ++i) {
                fv[(size_t)i] = facets[f * n + i];
                fn[(size_t)i] = has_normal[(size_t)f] ? facet_normals[f * n + i] : -1;
            }
            if (has_normal[(size_t)f] && n == 3) {
                // write_facets (src/obj/file/file_save.cpp:194-237): the order
                // follows the vertex normals
                double p[3][3];
                for (int k = 0; k < 3; ++k)
                    for (int i = 0; i < 3; ++i) p[k][i] = vertices[(int64_t)fv[(size_t)k] * 3 + i];
                double a[3], b[3];
                for (int i = 0; i < 3; ++i) {
                    a[i] = p[1][i] - p[0][i];
                    b[i] = p[2][i] - p[0][i];
                }
                const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                                     a[0] * b[1] - a[1] * b[0]};
                bool against = true;
                for (int k = 0; k < 3 && against; ++k) {
                    const float* q = normals + (int64_t)fn[(size_t)k] * 3;
                    double d = (double)q[0] * c[0];
                    d = std::fma((double)q[1], c[1], d);
                    d = std::fma((double)q[2], c[2], d);
                    against = d < 0;
                }
                if (against) {
                    std::swap(fv[1], fv[2]);
                    std::swap(fn[1], fn[2]);
                }
            }
            std::fputs("f", file);
            for (int i = 0; i < n; ++i) {
                if (has_normal[(size_t)f])
                    std::fprintf(file, " %ld//%ld", (long)fv[(size_t)i] + 1, (long)fn[(size_t)i] + 1);
                else
                    std::fprintf(file, " %ld", (long)fv[(size_t)i] + 1);
            }
            std::fputs("\n", file);
        }
        const int rc = std::fclose(file);
        file = nullptr;
        if (rc != 0) fail(std::string("Failed to write file ") + path);
        return NSP_OK;
    } catch (const ObjError& e) {
        if (file) std::fclose(file);
        g_obj_error = e.msg;
        return NSP_E_INVALID;
    } catch (const std::bad_alloc&) {
        if (file) std::fclose(file);
        g_obj_error = "host allocation failed";
        return NSP_E_NOMEM;
    }
}

}  // extern "C"
