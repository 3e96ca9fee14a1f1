// refconfig.cpp -- text loader for the reference's configuration headers
// (include/bcs_refconfig.h).  A small scanner, not a C++ parser: it reads the
// constructs those headers are made of -- `using Name = mp_list< ... >;`
// lists of Spring<...> / mp_float3<...> / mp_float<...> entries, the
// BloodCellDef<...> list, and the veinPositions / veinIndices arrays -- after
// removing comments.
#include "../../include/bcs_refconfig.h"
#include "bcs_internal.hpp"

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

using namespace bcs;

struct bcs_ref_config {
    struct Type {
        int32_t count = 0, ppc = 0;
        std::vector<int32_t> sa, sb;
        std::vector<float> sl, v;
    };
    std::vector<Type> types;
    std::vector<bcs_cell_type> ctypes;
    std::vector<float> pos, endc, endr;
    std::vector<uint32_t> idx;
};

namespace {

bool read_file(const char* path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// Comments removed (string literals do not occur in these headers).
std::string strip_comments(const std::string& s)
{
    std::string o;
    o.reserve(s.size());
    for (size_t i = 0; i < s.size();) {
        if (s.compare(i, 2, "//") == 0) {
            while (i < s.size() && s[i] != '\n') ++i;
        } else if (s.compare(i, 2, "/*") == 0) {
            const size_t e = s.find("*/", i + 2);
            i = e == std::string::npos ? s.size() : e + 2;
        } else {
            o += s[i++];
        }
    }
    return o;
}

// float(X) / 10^(p-1) as the reference's templates evaluate it: the integer
// power is converted to float and the quotient is rounded to float.
float fixed_point(long long x, int p)
{
    long long pw = 1;
    for (int k = 1; k < p; ++k) pw *= 10;
    return (float)x / (float)pw;
}

// Body of the template argument list that starts at s[open] == '<' (depth
// counted), or npos.
size_t match_angle(const std::string& s, size_t open)
{
    int depth = 0;
    for (size_t i = open; i < s.size(); ++i) {
        if (s[i] == '<') ++depth;
        else if (s[i] == '>' && --depth == 0) return i;
    }
    return std::string::npos;
}

std::vector<std::string> split_args(const std::string& body)
{
    std::vector<std::string> out;
    int depth = 0;
    std::string cur;
    for (char c : body) {
        if (c == '<') ++depth;
        if (c == '>') --depth;
        if (c == ',' && depth == 0) {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    out.push_back(cur);
    for (auto& a : out) {
        size_t b = a.find_first_not_of(" \t\r\n"), e = a.find_last_not_of(" \t\r\n");
        a = b == std::string::npos ? std::string() : a.substr(b, e - b + 1);
    }
    return out;
}

bool to_ll(const std::string& s, long long& v)
{
    if (s.empty()) return false;
    char* end = nullptr;
    v = std::strtoll(s.c_str(), &end, 10);
    return end && *end == 0;
}

// Every `using Name = mp_list< entries >;` as (entry template, integer args).
struct Entry {
    std::string name;
    std::vector<long long> args;
};
int read_lists(const std::string& text, const char* file, std::map<std::string, std::vector<Entry>>& lists)
{
    size_t at = 0;
    while ((at = text.find("using", at)) != std::string::npos) {
        size_t i = at + 5;
        at = i;
        if (i < text.size() && (std::isalnum((unsigned char)text[i]) || text[i] == '_')) continue;
        while (i < text.size() && std::isspace((unsigned char)text[i])) ++i;
        size_t n0 = i;
        while (i < text.size() && (std::isalnum((unsigned char)text[i]) || text[i] == '_')) ++i;
        const std::string lname = text.substr(n0, i - n0);
        while (i < text.size() && std::isspace((unsigned char)text[i])) ++i;
        if (lname.empty() || i >= text.size() || text[i] != '=') continue;
        ++i;
        while (i < text.size() && std::isspace((unsigned char)text[i])) ++i;
        if (text.compare(i, 7, "mp_list") != 0) continue;
        const size_t open = text.find('<', i);
        const size_t close = open == std::string::npos ? open : match_angle(text, open);
        if (close == std::string::npos)
            return fail(BCS_ERR_INVALID_ARGUMENT, std::string(file) + ": unterminated mp_list " + lname);
        std::vector<Entry> entries;
        for (const std::string& a : split_args(text.substr(open + 1, close - open - 1))) {
            if (a.empty()) continue;
            const size_t lt = a.find('<');
            Entry e;
            if (lt == std::string::npos || a.back() != '>') {
                e.name = a;  // a type name (e.g. a BloodCellDef argument)
            } else {
                e.name = a.substr(0, lt);
                for (const std::string& v : split_args(a.substr(lt + 1, a.size() - lt - 2))) {
                    long long x;
                    if (!to_ll(v, x)) {
                        e.args.clear();
                        e.name = a;
                        break;
                    }
                    e.args.push_back(x);
                }
            }
            entries.push_back(e);
        }
        lists[lname] = entries;
        at = close;
    }
    return BCS_OK;
}

int load_cells(const std::string& defs, const std::string& presets, const char* dpath, const char* ppath,
               bcs_ref_config& c)
{
    std::map<std::string, std::vector<Entry>> plist;
    if (int rc = read_lists(presets, ppath, plist)) return rc;
    // BloodCellDef<Count, ParticlesInCell, IndicesInCell, Color, Springs, Vertices, Indices, Normals>
    size_t at = 0;
    while ((at = defs.find("BloodCellDef", at)) != std::string::npos) {
        const size_t open = defs.find_first_not_of(" \t\r\n", at + 12);
        if (open == std::string::npos || defs[open] != '<') {
            at += 12;
            continue;
        }
        const size_t close = match_angle(defs, open);
        if (close == std::string::npos) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(dpath) + ": unterminated BloodCellDef");
        const std::vector<std::string> a = split_args(defs.substr(open + 1, close - open - 1));
        at = close;
        long long count, ppc;
        if (a.size() != 8 || !to_ll(a[0], count) || !to_ll(a[1], ppc))
            return fail(BCS_ERR_INVALID_ARGUMENT, std::string(dpath) + ": BloodCellDef needs 8 arguments");
        auto lookup = [&](const std::string& ref, std::vector<Entry>*& out) {
            std::string n = ref;
            const size_t sc = n.rfind("::");
            if (sc != std::string::npos) n = n.substr(sc + 2);
            auto it = plist.find(n);
            if (it == plist.end()) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(ppath) + ": no list " + n);
            out = &it->second;
            return BCS_OK;
        };
        std::vector<Entry>*springs, *verts;
        if (int rc = lookup(a[4], springs)) return rc;
        if (int rc = lookup(a[5], verts)) return rc;
        bcs_ref_config::Type t;
        t.count = (int32_t)count;
        t.ppc = (int32_t)ppc;
        for (const Entry& e : *springs) {
            if (e.name != "Spring" || (e.args.size() != 3 && e.args.size() != 4))
                return fail(BCS_ERR_INVALID_ARGUMENT, std::string(ppath) + ": bad spring entry in " + a[4]);
            if (e.args[0] < 0 || e.args[0] >= ppc || e.args[1] < 0 || e.args[1] >= ppc)
                return fail(BCS_ERR_INVALID_ARGUMENT, std::string(ppath) + ": spring vertex out of range in " + a[4]);
            t.sa.push_back((int32_t)e.args[0]);
            t.sb.push_back((int32_t)e.args[1]);
            t.sl.push_back(fixed_point(e.args[2], e.args.size() == 4 ? (int)e.args[3] : 7));
        }
        for (const Entry& e : *verts) {
            if (e.name != "mp_float3" || (e.args.size() != 3 && e.args.size() != 4))
                return fail(BCS_ERR_INVALID_ARGUMENT, std::string(ppath) + ": bad vertex entry in " + a[5]);
            const int p = e.args.size() == 4 ? (int)e.args[3] : 7;
            for (int d = 0; d < 3; ++d) t.v.push_back(fixed_point(e.args[d], p));
        }
        if ((long long)t.v.size() != 3 * ppc)
            return fail(BCS_ERR_INVALID_ARGUMENT, std::string(ppath) + ": " + a[5] + " does not hold ParticlesInCell vertices");
        c.types.push_back(std::move(t));
    }
    if (c.types.empty()) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(dpath) + ": no BloodCellDef found");
    return BCS_OK;
}

// Body between the '{' after `name` and its '}'.
bool array_body(const std::string& s, const char* name, std::string& body)
{
    const size_t n = s.find(name);
    if (n == std::string::npos) return false;
    const size_t b = s.find('{', n);
    const size_t e = b == std::string::npos ? b : s.find('}', b);
    if (e == std::string::npos) return false;
    body = s.substr(b + 1, e - b - 1);
    return true;
}

int load_vein(const std::string& v, const char* path, bcs_ref_config& c)
{
    std::string body;
    if (!array_body(v, "veinPositions", body)) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": no veinPositions");
    for (size_t at = 0; (at = body.find("static_cast<float>(", at)) != std::string::npos;) {
        at += 19;
        char* end = nullptr;
        const double d = std::strtod(body.c_str() + at, &end);
        if (!end || *end != ')') return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": bad vein position literal");
        c.pos.push_back((float)d);
    }
    if (c.pos.empty() || c.pos.size() % 3) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": vein positions not in triples");
    if (!array_body(v, "veinIndices", body)) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": no veinIndices");
    const uint64_t nv = c.pos.size() / 3;
    for (const std::string& a : split_args(body)) {
        if (a.empty()) continue;
        long long x;
        if (!to_ll(a, x) || x < 0 || (uint64_t)x >= nv)
            return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": bad vein index " + a);
        c.idx.push_back((uint32_t)x);
    }
    if (c.idx.empty() || c.idx.size() % 3) return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": vein indices not in triples");
    std::map<std::string, std::vector<Entry>> lists;
    if (int rc = read_lists(v, path, lists)) return rc;
    auto ce = lists.find("VeinEndingCenters");
    auto re = lists.find("VeinEndingRadii");
    if (ce != lists.end() && re != lists.end()) {
        if (ce->second.size() != re->second.size())
            return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": ending centres and radii differ in number");
        for (const Entry& e : ce->second) {
            if (e.name != "mp_float3" || (e.args.size() != 3 && e.args.size() != 4))
                return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": bad VeinEndingCenters entry");
            const int p = e.args.size() == 4 ? (int)e.args[3] : 7;
            for (int d = 0; d < 3; ++d) c.endc.push_back(fixed_point(e.args[d], p));
        }
        for (const Entry& e : re->second) {
            if (e.name != "mp_float" || (e.args.size() != 1 && e.args.size() != 2))
                return fail(BCS_ERR_INVALID_ARGUMENT, std::string(path) + ": bad VeinEndingRadii entry");
            c.endr.push_back(fixed_point(e.args[0], e.args.size() == 2 ? (int)e.args[1] : 7));
        }
    }
    return BCS_OK;
}

}  // namespace

extern "C" int bcs_ref_config_load(const char* definitions_path, const char* presets_path, const char* vein_path,
                                   bcs_ref_config** out)
{
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ref_config_load: null output");
    *out = nullptr;
    if (definitions_path && !presets_path)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ref_config_load: definitions need the presets file");
    std::unique_ptr<bcs_ref_config> c(new bcs_ref_config());
    if (definitions_path) {
        std::string d, p;
        if (!read_file(definitions_path, d)) return fail(BCS_ERR_INVALID_ARGUMENT, std::string("cannot read ") + definitions_path);
        if (!read_file(presets_path, p)) return fail(BCS_ERR_INVALID_ARGUMENT, std::string("cannot read ") + presets_path);
        if (int rc = load_cells(strip_comments(d), strip_comments(p), definitions_path, presets_path, *c)) return rc;
    }
    if (vein_path) {
        std::string v;
        if (!read_file(vein_path, v)) return fail(BCS_ERR_INVALID_ARGUMENT, std::string("cannot read ") + vein_path);
        if (int rc = load_vein(strip_comments(v), vein_path, *c)) return rc;
    }
    for (const auto& t : c->types) {
        bcs_cell_type ct{};
        ct.count = t.count;
        ct.particles_in_cell = t.ppc;
        ct.spring_count = (int32_t)t.sl.size();
        ct.spring_start = t.sa.data();
        ct.spring_end = t.sb.data();
        ct.spring_length = t.sl.data();
        ct.vertices = t.v.data();
        ct.contact_radius = 0.0f;
        c->ctypes.push_back(ct);
    }
    *out = c.release();
    return BCS_OK;
}

extern "C" int bcs_ref_config_get(bcs_ref_config* cfg, int32_t* n_types, const bcs_cell_type** types,
                                  bcs_vein_mesh* mesh)
{
    if (!cfg) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ref_config_get: null config");
    if (n_types) *n_types = (int32_t)cfg->ctypes.size();
    if (types) *types = cfg->ctypes.empty() ? nullptr : cfg->ctypes.data();
    if (mesh) {
        *mesh = bcs_vein_mesh{};
        mesh->vertex_count = (int32_t)(cfg->pos.size() / 3);
        mesh->positions = cfg->pos.empty() ? nullptr : cfg->pos.data();
        mesh->triangle_count = (int32_t)(cfg->idx.size() / 3);
        mesh->indices = cfg->idx.empty() ? nullptr : cfg->idx.data();
        mesh->ending_count = (int32_t)cfg->endr.size();
        mesh->ending_centers = cfg->endc.empty() ? nullptr : cfg->endc.data();
        mesh->ending_radii = cfg->endr.empty() ? nullptr : cfg->endr.data();
    }
    return BCS_OK;
}

extern "C" int bcs_ref_config_destroy(bcs_ref_config* cfg)
{
    delete cfg;
    return BCS_OK;
}
