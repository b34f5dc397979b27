// HDF5 file-format reader and writer for the drivers' input and output files (the subset the
// reference's files use), behind the reference's helper functions of src/util.c:56-236 and the
// H5Fopen / H5Fcreate / H5Fclose calls of its drivers.
//
// Format facts used here (HDF5 file format specification, version 0 superblock family):
//   superblock v0/v1  -> root group symbol-table entry -> root object header (version 1)
//   group             =  symbol table message (B-tree v1 of type 0 + local heap of names)
//   B-tree v1 node    =  "TREE", type, level, entries, siblings, key0 child0 key1 ... keyN
//                        (level 0 children are symbol table nodes "SNOD")
//   symbol table node =  "SNOD", version 1, count, entries (name offset, object header address,
//                        cache type, scratch)
//   local heap        =  "HEAP", version 0, data size, free-list head, data address; free blocks
//                        are (next offset, size) pairs and the list ends with offset 1
//   object header v1  =  version 1, message count, reference count, header size, 8-byte aligned
//                        messages (type u16, size u16, flags u8, 3 reserved, data)
//   messages          =  0x01 dataspace, 0x03 datatype, 0x05 fill value, 0x08 layout,
//                        0x0C attribute, 0x10 continuation, 0x11 symbol table (others skipped)
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/util.h"

namespace {

// ---- types -----------------------------------------------------------------------------------
enum Kind { kFloat, kSigned, kUnsigned };
struct Type {
    Kind kind;
    int size;
};

bool type_of_id(hid_t id, Type* t)
{
    switch (id - MG_H5T_ID(0)) {
        case 1: *t = {kFloat, 8}; return true;
        case 2: *t = {kFloat, 4}; return true;
        case 3: *t = {kSigned, 4}; return true;
        case 4: *t = {kSigned, 8}; return true;
        case 5: *t = {kUnsigned, 1}; return true;
        case 6: *t = {kSigned, 1}; return true;
        case 7: *t = {kSigned, 2}; return true;
        case 8: *t = {kUnsigned, 4}; return true;
        case 9: *t = {kUnsigned, 8}; return true;
    }
    return false;
}

void report(const std::string& msg) { std::fprintf(stderr, "rqcopt_mg hdf5: %s\n", msg.c_str()); }

// element conversion between the supported types (integers are range-checked; floats to integers
// are not converted, as HDF5 does not convert them by default either)
bool convert(const unsigned char* src, Type st, unsigned char* dst, Type dt, size_t n, std::string* why)
{
    if (st.kind == dt.kind && st.size == dt.size) {
        std::memcpy(dst, src, n * (size_t)st.size);
        return true;
    }
    if (st.kind == kFloat && dt.kind != kFloat) {
        *why = "no conversion from a floating-point to an integer type";
        return false;
    }
    for (size_t i = 0; i < n; i++) {
        const unsigned char* s = src + i * st.size;
        unsigned char* d = dst + i * dt.size;
        if (st.kind == kFloat) {
            double v;
            if (st.size == 8) std::memcpy(&v, s, 8);
            else {
                float f;
                std::memcpy(&f, s, 4);
                v = f;
            }
            if (dt.size == 8) std::memcpy(d, &v, 8);
            else {
                const float f = (float)v;
                std::memcpy(d, &f, 4);
            }
            continue;
        }
        // integer source: sign- or zero-extend to 64 bits
        uint64_t u = 0;
        std::memcpy(&u, s, st.size);
        bool neg = false;
        if (st.kind == kSigned && st.size < 8 && (u >> (8 * st.size - 1)) & 1) u |= ~0ull << (8 * st.size);
        if (st.kind == kSigned) neg = (int64_t)u < 0;
        if (dt.kind == kFloat) {
            const double v = st.kind == kSigned ? (double)(int64_t)u : (double)u;
            if (dt.size == 8) std::memcpy(d, &v, 8);
            else {
                const float f = (float)v;
                std::memcpy(d, &f, 4);
            }
            continue;
        }
        bool ok;
        if (dt.kind == kSigned) {
            const int64_t lo = dt.size == 8 ? INT64_MIN : -(int64_t)(1ull << (8 * dt.size - 1));
            const int64_t hi = dt.size == 8 ? INT64_MAX : (int64_t)((1ull << (8 * dt.size - 1)) - 1);
            ok = st.kind == kSigned ? ((int64_t)u >= lo && (int64_t)u <= hi) : (u <= (uint64_t)hi);
        } else {
            const uint64_t hi = dt.size == 8 ? UINT64_MAX : (1ull << (8 * dt.size)) - 1;
            ok = !neg && u <= hi;
        }
        if (!ok) {
            *why = "integer value out of range of the memory type";
            return false;
        }
        std::memcpy(d, &u, dt.size);
    }
    return true;
}

// ---- reading ---------------------------------------------------------------------------------
struct Object {
    std::vector<uint64_t> dims;
    bool scalar = false;
    Type type{kFloat, 0};
    bool has_type = false, has_space = false, has_layout = false;
    int layout = -1;  // 0 compact, 1 contiguous
    uint64_t addr = 0, size = 0;
    std::vector<unsigned char> compact;
};

struct Reader {
    std::vector<unsigned char> b;
    int O = 8, Lz = 8;
    std::string err;

    bool fail(const std::string& m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool have(uint64_t at, uint64_t n) const { return at <= b.size() && n <= b.size() - at; }
    uint64_t uN(uint64_t at, int n) const
    {
        uint64_t v = 0;
        for (int i = 0; i < n; i++) v |= (uint64_t)b[at + i] << (8 * i);
        return v;
    }

    // datatype message -> Type
    bool parse_type(uint64_t at, uint64_t n, Type* t)
    {
        if (n < 8 || !have(at, n)) return fail("truncated datatype message");
        const int cls = b[at] & 0x0f;
        const int bits0 = b[at + 1];
        const int size = (int)uN(at + 4, 4);
        if (cls == 0) {  // fixed point
            if (bits0 & 1) return fail("big-endian integers are not supported");
            if (size != 1 && size != 2 && size != 4 && size != 8) return fail("unsupported integer size");
            *t = {(bits0 & 8) ? kSigned : kUnsigned, size};
            return true;
        }
        if (cls == 1) {  // floating point
            if ((bits0 & 1) || (bits0 & 0x40)) return fail("big-endian or VAX floats are not supported");
            if (size != 4 && size != 8) return fail("unsupported float size");
            *t = {kFloat, size};
            return true;
        }
        return fail("unsupported datatype class " + std::to_string(cls) + " (only numbers)");
    }

    bool parse_space(uint64_t at, uint64_t n, Object* o)
    {
        if (n < 4 || !have(at, n)) return fail("truncated dataspace message");
        const int ver = b[at], nd = b[at + 1];
        uint64_t p;
        if (ver == 1) {
            p = at + 8;
            o->scalar = nd == 0;
        } else if (ver == 2) {
            p = at + 4;
            if (b[at + 3] == 2) return fail("null dataspace");
            o->scalar = b[at + 3] == 0;
        } else {
            return fail("unsupported dataspace version " + std::to_string(ver));
        }
        if (!have(p, (uint64_t)nd * Lz)) return fail("truncated dataspace dimensions");
        o->dims.resize(nd);
        for (int i = 0; i < nd; i++) o->dims[i] = uN(p + (uint64_t)i * Lz, Lz);
        o->has_space = true;
        return true;
    }

    bool parse_layout(uint64_t at, uint64_t n, Object* o)
    {
        if (n < 2 || !have(at, n)) return fail("truncated layout message");
        const int ver = b[at];
        if (ver == 3 || ver == 4) {
            const int cls = b[at + 1];
            if (cls == 0) {
                if (!have(at + 2, 2)) return fail("truncated compact layout");
                const uint64_t sz = uN(at + 2, 2);
                if (!have(at + 4, sz)) return fail("truncated compact data");
                o->compact.assign(b.begin() + at + 4, b.begin() + at + 4 + sz);
                o->layout = 0;
            } else if (cls == 1) {
                if (!have(at + 2, O + Lz)) return fail("truncated contiguous layout");
                o->addr = uN(at + 2, O);
                o->size = uN(at + 2 + O, Lz);
                o->layout = 1;
            } else {
                return fail("chunked or virtual datasets are not supported");
            }
        } else if (ver == 1 || ver == 2) {
            const int nd = b[at + 1], cls = b[at + 2];
            uint64_t p = at + 8;
            if (cls == 1) {
                o->addr = uN(p, O);
                p += O;
                o->layout = 1;
                o->size = UINT64_MAX;  // size from the dataspace and type
            } else if (cls == 0) {
                o->layout = 0;
            } else {
                return fail("chunked datasets are not supported");
            }
            p += 4ull * nd;
            if (cls == 0) {
                if (!have(p, 4)) return fail("truncated compact layout");
                const uint64_t sz = uN(p, 4);
                if (!have(p + 4, sz)) return fail("truncated compact data");
                o->compact.assign(b.begin() + p + 4, b.begin() + p + 4 + sz);
            }
        } else {
            return fail("unsupported layout version " + std::to_string(ver));
        }
        o->has_layout = true;
        return true;
    }

    // calls f(type, data address, size) for every message of a version-1 object header
    template <typename F>
    bool messages(uint64_t oh, F&& f)
    {
        if (!have(oh, 16)) return fail("object header out of the file");
        if (b[oh] != 1) {
            if (have(oh, 4) && std::memcmp(&b[oh], "OHDR", 4) == 0)
                return fail("version-2 object headers are not supported (file written with a newer format)");
            return fail("unsupported object header version " + std::to_string(b[oh]));
        }
        const uint64_t nmsg = uN(oh + 2, 2);
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{oh + 16, uN(oh + 8, 4)}};
        uint64_t seen = 0;
        for (size_t k = 0; k < blocks.size() && k < 4096; k++) {
            uint64_t p = blocks[k].first;
            const uint64_t end = p + blocks[k].second;
            if (!have(blocks[k].first, blocks[k].second)) return fail("object header block out of the file");
            while (p + 8 <= end && seen < nmsg) {
                const int type = (int)uN(p, 2);
                const uint64_t size = uN(p + 2, 2);
                const int flags = b[p + 4];
                const uint64_t d = p + 8;
                if (d + size > end) return fail("object header message overruns its block");
                seen++;
                if (type == 0x10) {
                    if (size < (uint64_t)(O + Lz)) return fail("truncated continuation message");
                    blocks.push_back({uN(d, O), uN(d + O, Lz)});
                } else if (type != 0) {
                    if (flags & 2) return fail("shared object header messages are not supported");
                    if (!f(type, d, size)) return false;
                }
                p = d + size;
            }
        }
        return true;
    }

    bool object(uint64_t oh, Object* o)
    {
        bool ok = messages(oh, [&](int type, uint64_t d, uint64_t n) {
            if (type == 0x01) return parse_space(d, n, o);
            if (type == 0x03) {
                o->has_type = true;
                return parse_type(d, n, &o->type);
            }
            if (type == 0x08) return parse_layout(d, n, o);
            return true;
        });
        if (!ok) return false;
        if (!o->has_space || !o->has_type || !o->has_layout) return fail("object is not a dataset");
        return true;
    }

    // attribute message -> name, Object (compact data)
    bool attribute(uint64_t at, uint64_t n, std::string* name, Object* o)
    {
        if (n < 8 || !have(at, n)) return fail("truncated attribute message");
        const int ver = b[at];
        const uint64_t nl = uN(at + 2, 2), tl = uN(at + 4, 2), sl = uN(at + 6, 2);
        uint64_t p = at + 8;
        auto pad = [&](uint64_t x) { return ver == 1 ? (x + 7) / 8 * 8 : x; };
        if (ver == 3) p++;  // name character set
        else if (ver != 1 && ver != 2) return fail("unsupported attribute message version");
        if (!have(p, pad(nl) + pad(tl) + pad(sl))) return fail("truncated attribute message");
        name->assign((const char*)&b[p], nl ? nl - 1 : 0);
        p += pad(nl);
        if (!parse_type(p, tl, &o->type)) return false;
        p += pad(tl);
        if (!parse_space(p, sl, o)) return false;
        p += pad(sl);
        uint64_t cnt = 1;
        for (uint64_t d : o->dims) cnt *= d;
        const uint64_t bytes = cnt * (uint64_t)o->type.size;
        if (p + bytes > at + n || !have(p, bytes)) return fail("truncated attribute data");
        o->compact.assign(b.begin() + p, b.begin() + p + bytes);
        o->layout = 0;
        o->has_type = o->has_layout = true;
        return true;
    }

    // superblock -> address of the root group's object header
    bool superblock(uint64_t* root_oh)
    {
        if (b.size() < 8 || std::memcmp(b.data(), "\x89HDF\r\n\x1a\n", 8) != 0) return fail("not an HDF5 file");
        if (!have(0, 24)) return fail("truncated superblock");
        const int sbv = b[8];
        if (sbv > 1) return fail("superblock version " + std::to_string(sbv) + " is not supported (only 0 and 1)");
        O = b[13];
        Lz = b[14];
        if (O != 8 || Lz != 8) return fail("only 8-byte offsets and lengths are supported");
        const uint64_t p = sbv == 0 ? 24 : 28;
        if (!have(p, 4ull * O)) return fail("truncated superblock");
        if (uN(p, O) != 0) return fail("files with a user block are not supported");
        const uint64_t ent = p + 4ull * O;  // root symbol-table entry
        if (!have(ent, 2ull * O + 24)) return fail("truncated superblock");
        *root_oh = uN(ent + O, O);
        return true;
    }

    // a group's members (name -> object header address) and attributes; `is_group` is false
    // for an object without a symbol table message (a dataset)
    bool group(uint64_t oh, std::map<std::string, uint64_t>* links, std::map<std::string, Object>* attrs,
               bool* is_group)
    {
        uint64_t btree = UINT64_MAX, heap = UINT64_MAX;
        bool ok = messages(oh, [&](int type, uint64_t d, uint64_t n) {
            if (type == 0x11) {
                if (n < 2ull * O) return fail("truncated symbol table message");
                btree = uN(d, O);
                heap = uN(d + O, O);
            } else if (type == 0x0C && attrs) {
                std::string name;
                Object a;
                if (!attribute(d, n, &name, &a)) return false;
                (*attrs)[name] = std::move(a);
            } else if (type == 0x06 || type == 0x02 || type == 0x0A) {
                return fail("new-style (link message) groups are not supported");
            }
            return true;
        });
        if (!ok) return false;
        *is_group = btree != UINT64_MAX;
        if (!*is_group) return true;
        if (!have(heap, 32) || std::memcmp(&b[heap], "HEAP", 4) != 0) return fail("bad local heap");
        const uint64_t hsize = uN(heap + 8, Lz), hdata = uN(heap + 24, O);
        if (!have(hdata, hsize)) return fail("local heap data out of the file");
        auto heap_name = [&](uint64_t off, std::string* s) {
            if (off >= hsize) return fail("name offset outside the local heap");
            const char* c = (const char*)&b[hdata + off];
            const size_t len = strnlen(c, hsize - off);
            if (len == hsize - off) return fail("unterminated name in the local heap");
            s->assign(c, len);
            return true;
        };
        // walk the B-tree (depth-bounded)
        std::vector<std::pair<uint64_t, int>> todo{{btree, 0}};
        size_t visited = 0;
        while (!todo.empty()) {
            const auto [node, depth] = todo.back();
            todo.pop_back();
            if (++visited > 100000 || depth > 64) return fail("group B-tree too deep or cyclic");
            if (!have(node, 8) || std::memcmp(&b[node], "TREE", 4) != 0) return fail("bad B-tree node");
            if (b[node + 4] != 0) return fail("B-tree node is not a group node");
            const int level = b[node + 5];
            const uint64_t used = uN(node + 6, 2);
            const uint64_t base = node + 8 + 2ull * O;
            if (!have(base, (used + 1) * Lz + used * O)) return fail("truncated B-tree node");
            for (uint64_t i = 0; i < used; i++) {
                const uint64_t child = uN(base + Lz + i * (Lz + O), O);
                if (level > 0) {
                    todo.push_back({child, depth + 1});
                    continue;
                }
                if (!have(child, 8) || std::memcmp(&b[child], "SNOD", 4) != 0) return fail("bad symbol table node");
                const uint64_t cnt = uN(child + 6, 2);
                const uint64_t esz = 2ull * O + 24;
                if (!have(child + 8, cnt * esz)) return fail("truncated symbol table node");
                for (uint64_t e = 0; e < cnt; e++) {
                    const uint64_t q = child + 8 + e * esz;
                    std::string name;
                    if (!heap_name(uN(q, O), &name)) return false;
                    (*links)[name] = uN(q + O, O);
                }
            }
        }
        return true;
    }

    bool data(const Object& o, std::vector<unsigned char>* out)
    {
        uint64_t cnt = 1;
        for (uint64_t d : o.dims) {
            if (d && cnt > UINT64_MAX / d) return fail("dataset too large");
            cnt *= d;
        }
        const uint64_t bytes = cnt * (uint64_t)o.type.size;
        if (o.layout == 0) {
            if (o.compact.size() < bytes) return fail("compact data shorter than the dataspace");
            out->assign(o.compact.begin(), o.compact.begin() + bytes);
            return true;
        }
        if (o.size != UINT64_MAX && o.size < bytes) return fail("contiguous storage shorter than the dataspace");
        if (o.addr == UINT64_MAX) {  // never written: the default fill value (zeros)
            out->assign(bytes, 0);
            return true;
        }
        if (!have(o.addr, bytes)) return fail("dataset storage out of the file");
        out->assign(b.begin() + o.addr, b.begin() + o.addr + bytes);
        return true;
    }
};

// ---- writing ---------------------------------------------------------------------------------
struct Pending {
    std::string name;
    std::vector<uint64_t> dims;  // empty: scalar
    Type type;
    std::vector<unsigned char> bytes;
};

struct File {
    bool writing = false;
    std::string path;
    Reader rd;  // reading: the whole file, its root group and its attributes
    uint64_t root_oh = 0;
    std::map<std::string, uint64_t> links;
    std::map<std::string, Object> attrs;

    // object header address of a path of group members ("a", "g/a", "/g/a"), or UINT64_MAX
    uint64_t resolve(const std::string& path)
    {
        std::vector<std::string> parts;
        size_t i = 0;
        while (i <= path.size()) {
            const size_t j = std::min(path.find('/', i), path.size());
            if (j > i) parts.push_back(path.substr(i, j - i));
            i = j + 1;
        }
        if (parts.empty()) return UINT64_MAX;
        std::map<std::string, uint64_t> members = links;
        for (size_t k = 0; k < parts.size(); k++) {
            auto it = members.find(parts[k]);
            if (it == members.end()) return UINT64_MAX;
            if (k + 1 == parts.size()) return it->second;
            std::map<std::string, uint64_t> next;
            bool is_group = false;
            if (!rd.group(it->second, &next, nullptr, &is_group) || !is_group) return UINT64_MAX;
            members.swap(next);
        }
        return UINT64_MAX;
    }
    std::vector<Pending> dsets, wattrs;  // writing
};

std::mutex g_mu;
std::map<hid_t, std::unique_ptr<File>> g_files;
hid_t g_next = 1;

File* lookup(hid_t h)
{
    auto it = g_files.find(h);
    return it == g_files.end() ? nullptr : it->second.get();
}

struct Out {
    std::vector<unsigned char> b;
    uint64_t pos() const { return b.size(); }
    void u(uint64_t v, int n)
    {
        for (int i = 0; i < n; i++) b.push_back((unsigned char)(v >> (8 * i)));
    }
    void raw(const void* p, size_t n) { b.insert(b.end(), (const unsigned char*)p, (const unsigned char*)p + n); }
    void pad8()
    {
        while (b.size() % 8) b.push_back(0);
    }
    void put(uint64_t at, uint64_t v, int n)
    {
        for (int i = 0; i < n; i++) b[at + i] = (unsigned char)(v >> (8 * i));
    }
};

std::vector<unsigned char> type_msg(Type t)
{
    Out m;
    if (t.kind == kFloat) {
        // class 1 version 1; little endian, implied mantissa msb, sign bit at the top
        m.u(0x11, 1);
        m.u(0x20, 1);
        m.u(8 * t.size - 1, 1);
        m.u(0, 1);
        m.u(t.size, 4);
        m.u(0, 2);                                 // bit offset
        m.u(8 * t.size, 2);                        // precision
        m.u(t.size == 8 ? 52 : 23, 1);             // exponent location
        m.u(t.size == 8 ? 11 : 8, 1);              // exponent size
        m.u(0, 1);                                 // mantissa location
        m.u(t.size == 8 ? 52 : 23, 1);             // mantissa size
        m.u(t.size == 8 ? 1023 : 127, 4);          // exponent bias
    } else {
        m.u(0x10, 1);
        m.u(t.kind == kSigned ? 0x08 : 0x00, 1);
        m.u(0, 1);
        m.u(0, 1);
        m.u(t.size, 4);
        m.u(0, 2);
        m.u(8 * t.size, 2);
    }
    return m.b;
}

std::vector<unsigned char> space_msg(const std::vector<uint64_t>& dims)
{
    Out m;
    m.u(1, 1);                          // version 1
    m.u(dims.size(), 1);                // 0: scalar
    m.u(dims.empty() ? 0 : 1, 1);       // maximum dimensions present
    m.u(0, 5);
    for (uint64_t d : dims) m.u(d, 8);
    for (uint64_t d : dims) m.u(d, 8);  // fixed size: max = current
    return m.b;
}

void message(Out* o, int type, const std::vector<unsigned char>& data, int flags = 0)
{
    const size_t n = (data.size() + 7) / 8 * 8;
    o->u(type, 2);
    o->u(n, 2);
    o->u(flags, 1);
    o->u(0, 3);
    o->raw(data.data(), data.size());
    for (size_t i = data.size(); i < n; i++) o->b.push_back(0);
}

std::vector<unsigned char> attribute_msg(const Pending& a)
{
    Out m;
    const std::vector<unsigned char> t = type_msg(a.type), s = space_msg({});
    m.u(1, 1);
    m.u(0, 1);
    m.u(a.name.size() + 1, 2);
    m.u(t.size(), 2);
    m.u(s.size(), 2);
    m.raw(a.name.c_str(), a.name.size() + 1);
    m.pad8();
    m.raw(t.data(), t.size());
    m.pad8();
    m.raw(s.data(), s.size());
    m.pad8();
    m.raw(a.bytes.data(), a.bytes.size());
    return m.b;
}

// version-1 object header holding the given messages (type, data)
uint64_t object_header(Out* o, const std::vector<std::pair<int, std::vector<unsigned char>>>& msgs)
{
    Out body;
    for (const auto& m : msgs) message(&body, m.first, m.second, m.first == 0x03 || m.first == 0x05 ? 1 : 0);
    const uint64_t at = o->pos();
    o->u(1, 1);
    o->u(0, 1);
    o->u(msgs.size(), 2);
    o->u(1, 4);                 // reference count
    o->u(body.b.size(), 4);     // header size
    o->u(0, 4);                 // alignment to 8 bytes
    o->raw(body.b.data(), body.b.size());
    return at;
}

constexpr int kLeafK = 4, kInternalK = 16;  // the library's defaults

bool write_file(const File& f, std::string* why)
{
    std::vector<const Pending*> ds;
    for (const Pending& p : f.dsets) ds.push_back(&p);
    std::sort(ds.begin(), ds.end(), [](const Pending* a, const Pending* b) { return a->name < b->name; });
    const size_t per = 2 * kLeafK;
    const size_t nsnod = std::max<size_t>(1, (ds.size() + per - 1) / per);
    if (nsnod > 2 * kInternalK) {
        *why = "too many datasets for one group node (at most " + std::to_string(2 * kInternalK * per) + ")";
        return false;
    }
    // local heap: "" at 0, then the names, 8-byte aligned; a trailing free block
    Out heap;
    heap.u(0, 8);
    std::vector<uint64_t> name_off(ds.size());
    for (size_t i = 0; i < ds.size(); i++) {
        name_off[i] = heap.pos();
        heap.raw(ds[i]->name.c_str(), ds[i]->name.size() + 1);
        heap.pad8();
    }
    const uint64_t free_off = heap.pos();
    heap.u(1, 8);   // end of the free list
    heap.u(16, 8);  // size of this free block

    Out o;
    // superblock version 0
    o.raw("\x89HDF\r\n\x1a\n", 8);
    o.u(0, 1);  // superblock version
    o.u(0, 1);  // free-space storage version
    o.u(0, 1);  // root group symbol table entry version
    o.u(0, 1);
    o.u(0, 1);  // shared header message format version
    o.u(8, 1);  // size of offsets
    o.u(8, 1);  // size of lengths
    o.u(0, 1);
    o.u(kLeafK, 2);
    o.u(kInternalK, 2);
    o.u(0, 4);  // file consistency flags
    o.u(0, 8);                 // base address
    o.u(UINT64_MAX, 8);        // free-space info address
    const uint64_t eof_at = o.pos();
    o.u(0, 8);                 // end of file address (patched)
    o.u(UINT64_MAX, 8);        // driver information block
    // root group symbol table entry (patched)
    const uint64_t ent = o.pos();
    o.u(0, 8);   // link name offset
    o.u(0, 8);   // object header address
    o.u(1, 4);   // cache type 1: scratch holds the B-tree and heap addresses
    o.u(0, 4);
    o.u(0, 8);
    o.u(0, 8);

    // datasets: object header then data
    std::vector<uint64_t> ds_oh(ds.size());
    for (size_t i = 0; i < ds.size(); i++) {
        const Pending& p = *ds[i];
        Out layout;
        // layout version 3, contiguous: the address is patched once the data is placed
        layout.u(3, 1);
        layout.u(1, 1);
        layout.u(0, 8);
        layout.u(p.bytes.size(), 8);
        const std::vector<unsigned char> fill{2, 2, 2, 1, 0, 0, 0, 0};  // version 2, late allocation, default value
        o.pad8();
        const uint64_t oh = object_header(&o, {{0x01, space_msg(p.dims)}, {0x03, type_msg(p.type)}, {0x05, fill},
                                               {0x08, layout.b}});
        ds_oh[i] = oh;
        o.pad8();
        const uint64_t data_at = o.pos();
        o.raw(p.bytes.data(), p.bytes.size());
        // the layout message is the fourth: skip the three before it (8-byte prefix + size each)
        uint64_t q = oh + 16;
        for (int k = 0; k < 3; k++) q += 8 + ((uint64_t)o.b[q + 2] | ((uint64_t)o.b[q + 3] << 8));
        o.put(q + 8 + 2, data_at, 8);
    }

    // symbol table nodes
    o.pad8();
    std::vector<uint64_t> snod(nsnod);
    for (size_t s = 0; s < nsnod; s++) {
        snod[s] = o.pos();
        const size_t e0 = s * per, e1 = std::min(ds.size(), e0 + per);
        o.raw("SNOD", 4);
        o.u(1, 1);
        o.u(0, 1);
        o.u(e1 - e0, 2);
        for (size_t e = 0; e < per; e++) {
            if (e0 + e < e1) {
                o.u(name_off[e0 + e], 8);
                o.u(ds_oh[e0 + e], 8);
            } else {
                o.u(0, 8);
                o.u(0, 8);
            }
            o.u(0, 4);
            o.u(0, 4);
            o.u(0, 8);
            o.u(0, 8);
        }
    }
    // B-tree node (level 0) over the symbol table nodes; keys: "" then each node's last name
    const uint64_t btree = o.pos();
    o.raw("TREE", 4);
    o.u(0, 1);
    o.u(0, 1);
    o.u(nsnod, 2);
    o.u(UINT64_MAX, 8);
    o.u(UINT64_MAX, 8);
    o.u(0, 8);
    for (size_t s = 0; s < 2 * kInternalK; s++) {
        if (s < nsnod) {
            const size_t last = std::min(ds.size(), (s + 1) * per);
            o.u(snod[s], 8);
            o.u(last ? name_off[last - 1] : 0, 8);
        } else {
            o.u(0, 8);
            o.u(0, 8);
        }
    }
    // local heap
    const uint64_t heap_at = o.pos();
    o.raw("HEAP", 4);
    o.u(0, 1);
    o.u(0, 3);
    o.u(heap.pos(), 8);
    o.u(free_off, 8);
    o.u(heap_at + 32, 8);
    o.raw(heap.b.data(), heap.b.size());

    // root object header: symbol table message and the attributes
    std::vector<std::pair<int, std::vector<unsigned char>>> rm;
    Out st;
    st.u(btree, 8);
    st.u(heap_at, 8);
    rm.push_back({0x11, st.b});
    for (const Pending& a : f.wattrs) rm.push_back({0x0C, attribute_msg(a)});
    o.pad8();
    const uint64_t root_oh = object_header(&o, rm);
    o.pad8();

    o.put(ent + 8, root_oh, 8);
    o.put(ent + 24, btree, 8);
    o.put(ent + 32, heap_at, 8);
    o.put(eof_at, o.pos(), 8);

    FILE* fp = std::fopen(f.path.c_str(), "wb");
    if (!fp) {
        *why = "cannot open '" + f.path + "' for writing";
        return false;
    }
    const bool ok = std::fwrite(o.b.data(), 1, o.b.size(), fp) == o.b.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) {
        *why = "writing '" + f.path + "' failed";
        return false;
    }
    return true;
}

bool valid_name(const char* name) { return name && *name && !std::strchr(name, '/'); }

}  // namespace

extern "C" {

hid_t mg_h5_fopen(const char* name, unsigned flags)
{
    if (!name) return -1;
    if (flags & (H5F_ACC_RDWR | H5F_ACC_TRUNC)) {
        report("H5Fopen: only read-only access is supported");
        return -1;
    }
    auto f = std::make_unique<File>();
    FILE* fp = std::fopen(name, "rb");
    if (!fp) {
        report(std::string("H5Fopen: cannot open '") + name + "'");
        return -1;
    }
    std::fseek(fp, 0, SEEK_END);
    const long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    f->rd.b.resize(n > 0 ? (size_t)n : 0);
    const bool ok = n >= 0 && std::fread(f->rd.b.data(), 1, f->rd.b.size(), fp) == f->rd.b.size();
    std::fclose(fp);
    if (!ok) {
        report(std::string("H5Fopen: reading '") + name + "' failed");
        return -1;
    }
    bool is_group = false;
    if (!f->rd.superblock(&f->root_oh) || !f->rd.group(f->root_oh, &f->links, &f->attrs, &is_group) || !is_group) {
        if (f->rd.err.empty()) f->rd.err = "root group has no symbol table";
        report(std::string("H5Fopen: '") + name + "': " + f->rd.err);
        return -1;
    }
    f->path = name;
    std::lock_guard<std::mutex> lk(g_mu);
    const hid_t h = g_next++;
    g_files[h] = std::move(f);
    return h;
}

hid_t mg_h5_fcreate(const char* name, unsigned flags)
{
    if (!name) return -1;
    if (!(flags & H5F_ACC_TRUNC)) {
        if (FILE* fp = std::fopen(name, "rb")) {
            std::fclose(fp);
            report(std::string("H5Fcreate: '") + name + "' exists (use H5F_ACC_TRUNC)");
            return -1;
        }
    }
    // create the file now so that a bad path fails here, as with libhdf5
    FILE* fp = std::fopen(name, "wb");
    if (!fp) {
        report(std::string("H5Fcreate: cannot create '") + name + "'");
        return -1;
    }
    std::fclose(fp);
    auto f = std::make_unique<File>();
    f->writing = true;
    f->path = name;
    std::lock_guard<std::mutex> lk(g_mu);
    const hid_t h = g_next++;
    g_files[h] = std::move(f);
    return h;
}

herr_t mg_h5_fclose(hid_t file)
{
    std::unique_ptr<File> f;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_files.find(file);
        if (it == g_files.end()) {
            report("H5Fclose: invalid file handle");
            return -1;
        }
        f = std::move(it->second);
        g_files.erase(it);
    }
    if (f->writing) {
        std::string why;
        if (!write_file(*f, &why)) {
            report("H5Fclose: " + why);
            return -1;
        }
    }
    return 0;
}

static herr_t find_dataset(hid_t file, const char* name, const char* fn, Object* o, Reader** rd)
{
    File* f = lookup(file);
    if (!f || f->writing) {
        report(std::string(fn) + ": invalid or write-only file handle");
        return -1;
    }
    const uint64_t oh = f->resolve(name ? name : "");
    if (oh == UINT64_MAX) {
        report(std::string(fn) + ": no dataset '" + (name ? name : "") + "'");
        f->rd.err.clear();
        return -1;
    }
    if (!f->rd.object(oh, o)) {
        report(std::string(fn) + ": '" + name + "': " + f->rd.err);
        f->rd.err.clear();
        return -1;
    }
    *rd = &f->rd;
    return 0;
}

herr_t get_hdf5_dataset_ndims(hid_t file, const char* name, int* ndims)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Object o;
    Reader* rd;
    if (find_dataset(file, name, "get_hdf5_dataset_ndims", &o, &rd) < 0) return -1;
    *ndims = (int)o.dims.size();
    return 0;
}

herr_t get_hdf5_dataset_dims(hid_t file, const char* name, hsize_t* dims)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Object o;
    Reader* rd;
    if (find_dataset(file, name, "get_hdf5_dataset_dims", &o, &rd) < 0) return -1;
    for (size_t i = 0; i < o.dims.size(); i++) dims[i] = o.dims[i];
    return 0;
}

static herr_t read_object(Reader* rd, const Object& o, hid_t mem_type, void* data, const std::string& what)
{
    Type mt;
    if (!type_of_id(mem_type, &mt)) {
        report(what + ": unknown memory type");
        return -1;
    }
    std::vector<unsigned char> raw;
    if (!rd->data(o, &raw)) {
        report(what + ": " + rd->err);
        rd->err.clear();
        return -1;
    }
    std::string why;
    if (!convert(raw.data(), o.type, (unsigned char*)data, mt, raw.size() / o.type.size, &why)) {
        report(what + ": " + why);
        return -1;
    }
    return 0;
}

herr_t read_hdf5_dataset(hid_t file, const char* name, hid_t mem_type, void* data)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Object o;
    Reader* rd;
    if (find_dataset(file, name, "read_hdf5_dataset", &o, &rd) < 0) return -1;
    return read_object(rd, o, mem_type, data, std::string("read_hdf5_dataset: '") + name + "'");
}

herr_t read_hdf5_attribute(hid_t file, const char* name, hid_t mem_type, void* data)
{
    std::lock_guard<std::mutex> lk(g_mu);
    File* f = lookup(file);
    if (!f || f->writing) {
        report("read_hdf5_attribute: invalid or write-only file handle");
        return -1;
    }
    auto it = f->attrs.find(name ? name : "");
    if (it == f->attrs.end()) {
        report(std::string("read_hdf5_attribute: no attribute '") + (name ? name : "") + "'");
        return -1;
    }
    return read_object(&f->rd, it->second, mem_type, data, std::string("read_hdf5_attribute: '") + name + "'");
}

static herr_t add_pending(hid_t file, const char* name, int degree, const hsize_t* dims, hid_t store, hid_t input,
                          const void* data, bool attr)
{
    const char* fn = attr ? "write_hdf5_scalar_attribute" : "write_hdf5_dataset";
    std::lock_guard<std::mutex> lk(g_mu);
    File* f = lookup(file);
    if (!f || !f->writing) {
        report(std::string(fn) + ": invalid or read-only file handle");
        return -1;
    }
    if (!valid_name(name)) {
        report(std::string(fn) + ": invalid name");
        return -1;
    }
    Pending p;
    p.name = name;
    Type it;
    if (!type_of_id(store, &p.type) || !type_of_id(input, &it)) {
        report(std::string(fn) + ": unknown type");
        return -1;
    }
    if (degree < 0 || degree > 32 || (degree > 0 && !dims)) {
        report(std::string(fn) + ": invalid rank");
        return -1;
    }
    uint64_t cnt = 1;
    for (int i = 0; i < degree; i++) {
        p.dims.push_back(dims[i]);
        cnt *= dims[i];
    }
    std::vector<Pending>& list = attr ? f->wattrs : f->dsets;
    for (const Pending& q : list)
        if (q.name == p.name) {
            report(std::string(fn) + ": '" + name + "' exists");
            return -1;
        }
    if (cnt && !data) {
        report(std::string(fn) + ": no data");
        return -1;
    }
    p.bytes.resize(cnt * (uint64_t)p.type.size);
    std::string why;
    if (cnt && !convert((const unsigned char*)data, it, p.bytes.data(), p.type, cnt, &why)) {
        report(std::string(fn) + ": '" + name + "': " + why);
        return -1;
    }
    list.push_back(std::move(p));
    return 0;
}

herr_t write_hdf5_dataset(hid_t file, const char* name, int degree, const hsize_t dims[], hid_t mem_type_store,
                          hid_t mem_type_input, const void* data)
{
    return add_pending(file, name, degree, dims, mem_type_store, mem_type_input, data, false);
}

herr_t write_hdf5_scalar_attribute(hid_t file, const char* name, hid_t mem_type_store, hid_t mem_type_input,
                                   const void* data)
{
    return add_pending(file, name, 0, nullptr, mem_type_store, mem_type_input, data, true);
}

}  // extern "C"
