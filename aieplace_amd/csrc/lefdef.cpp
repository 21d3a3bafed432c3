// lefdef.cpp -- LEF/DEF reader and DEF writer (include/eplace_io.h).
//
// Replaces DataBase::readLEF / readDEF with the Limbo callbacks they feed
// (vck5000/host/src/DataBase.cpp:42-96, :318-424) and DataBase::writeDEF
// (:538-654).  A small tokenizer plus one routine per section; constructs the
// placer does not use are skipped by their LEF/DEF block structure, so the
// reader does not need a grammar for them.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <locale>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/eplace_io.h"
#include "ep_common.h"

using ep::fail;

namespace {

// ---------------------------------------------------------------- tokens

// Whitespace-separated tokens; '#' starts a comment; "..." is one token;
// unescaped '(' ')' ';' are tokens of their own (names escape them with '\').
struct Lexer {
    std::string text;
    size_t at = 0;
    int line = 1;
    std::string path;

    bool next(std::string &tok)
    {
        tok.clear();
        for (;;) {
            while (at < text.size() && std::isspace((unsigned char)text[at])) {
                if (text[at] == '\n') ++line;
                ++at;
            }
            if (at >= text.size()) return false;
            if (text[at] != '#') break;
            while (at < text.size() && text[at] != '\n') ++at;
        }
        const char c = text[at];
        if (c == '(' || c == ')' || c == ';') {
            tok.assign(1, c);
            ++at;
            return true;
        }
        if (c == '"') {
            const size_t e = text.find('"', at + 1);
            const size_t end = e == std::string::npos ? text.size() : e + 1;
            tok = text.substr(at, end - at);
            at = end;
            return true;
        }
        while (at < text.size()) {
            const char d = text[at];
            if (std::isspace((unsigned char)d) || d == '(' || d == ')' || d == ';') break;
            if (d == '\\' && at + 1 < text.size()) {  // escaped character stays in the name
                tok.push_back(d);
                tok.push_back(text[at + 1]);
                at += 2;
                continue;
            }
            tok.push_back(d);
            ++at;
        }
        return true;
    }
};

bool read_file(const std::string &path, std::string &out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool to_double(const std::string &s, double &v)
{
    if (s.empty()) return false;
    char *e = nullptr;
    v = std::strtod(s.c_str(), &e);
    return e && *e == '\0' && std::isfinite(v);
}

bool to_int64(const std::string &s, int64_t &v)
{
    double d;
    if (!to_double(s, d)) return false;
    v = (int64_t)std::llround(d);
    return true;
}

// ---------------------------------------------------------------- design data

struct Macro {
    double w = 0, h = 0, ox = 0, oy = 0;
    std::unordered_map<std::string, std::pair<double, double>> pins;  // centre, macro coordinates (um)
};

enum { kStatNone = 0, kStatPlaced, kStatUnplaced, kStatFixed, kStatCover };

struct Comp {
    std::string name, macro;
    int status = kStatNone;
    bool has_loc = false;
    int64_t x = 0, y = 0;
    std::string orient = "N";
};

struct IoPin {
    std::string name, net, dir, layer;
    int status = kStatNone;
    bool has_loc = false, has_box = false;
    int64_t x = 0, y = 0;
    int64_t box[4] = {0, 0, 0, 0};
    std::string orient = "N";
};

struct Conn {
    int io = 0;  // 1: IO pin (idx into pins), 0: component pin
    int idx = 0;
    std::string pin;
};

struct DNet {
    std::string name;
    std::vector<Conn> conn;
};

bool is_fixed(int status) { return status == kStatFixed || status == kStatCover; }

const char *status_word(int status)
{
    switch (status) {
    case kStatFixed: return "FIXED";
    case kStatCover: return "COVER";
    case kStatUnplaced: return "UNPLACED";
    default: return "PLACED";
    }
}

// Point (px, py) of a W x H macro, relative to its lower-left corner, under a
// DEF orientation; the oriented box's lower-left stays at the placement point.
bool orient_point(const std::string &o, double W, double H, double px, double py, double &qx, double &qy)
{
    if (o == "N") { qx = px; qy = py; }
    else if (o == "S") { qx = W - px; qy = H - py; }
    else if (o == "W") { qx = H - py; qy = px; }
    else if (o == "E") { qx = py; qy = W - px; }
    else if (o == "FN") { qx = W - px; qy = py; }
    else if (o == "FS") { qx = px; qy = H - py; }
    else if (o == "FW") { qx = py; qy = px; }
    else if (o == "FE") { qx = H - py; qy = W - px; }
    else return false;
    return true;
}

bool quarter_turn(const std::string &o) { return o == "W" || o == "E" || o == "FW" || o == "FE"; }

}  // namespace

struct ep_lefdef {
    std::string name = "design";
    int32_t units = 1000;
    int64_t die[4] = {0, 0, 0, 0};
    bool has_die = false;
    std::unordered_map<std::string, Macro> macros;
    std::vector<Comp> comps;
    std::vector<IoPin> pins;
    std::vector<DNet> nets;
    // netlist view
    int32_t n_mov = 0, n_fix = 0;
    std::vector<float> w, h, x, y, pin_dx, pin_dy;
    std::vector<int32_t> net_start, pin_cell;
    std::vector<int32_t> cell_of_comp, cell_of_pin;
    std::vector<int32_t> movable_comps;  // cell k < n_mov is component movable_comps[k]
};

namespace {

// ---------------------------------------------------------------- LEF

ep_status lef_error(const Lexer &L, const char *what)
{
    return fail(EP_EIO, "LEF %s:%d: %s", L.path.c_str(), L.line, what);
}

// Skips to the token `a` followed by `b` (b empty: just `a`); consumes both.
bool skip_to(Lexer &L, const std::string &a, const std::string &b = std::string())
{
    std::string t;
    while (L.next(t)) {
        if (t != a) continue;
        if (b.empty()) return true;
        const size_t save = L.at;
        const int line = L.line;
        std::string u;
        if (L.next(u) && u == b) return true;
        L.at = save;
        L.line = line;
    }
    return false;
}

bool skip_statement(Lexer &L)
{
    std::string t;
    while (L.next(t))
        if (t == ";") return true;
    return false;
}

// One PIN block: the bounding box of all its RECT / POLYGON shapes.
ep_status lef_pin(Lexer &L, const std::string &pname, Macro &m)
{
    double bx0 = 1e300, by0 = 1e300, bx1 = -1e300, by1 = -1e300;
    std::string t;
    while (L.next(t)) {
        if (t == "END") {
            const size_t save = L.at;
            const int line = L.line;
            std::string u;
            if (L.next(u) && u == pname) {
                if (bx1 >= bx0) m.pins[pname] = {0.5 * (bx0 + bx1), 0.5 * (by0 + by1)};
                return EP_OK;
            }
            L.at = save;
            L.line = line;
            continue;
        }
        if (t == "RECT" || t == "POLYGON") {
            std::vector<double> v;
            std::string u;
            while (L.next(u) && u != ";") {
                double d;
                if (u == "MASK" || u == "ITERATE") {  // MASK n: the colour; ITERATE: arrays are not used
                    L.next(u);
                    continue;
                }
                if (to_double(u, d)) v.push_back(d);
            }
            if (v.size() < 4 || (v.size() & 1)) return lef_error(L, "malformed RECT/POLYGON");
            for (size_t k = 0; k + 1 < v.size(); k += 2) {
                bx0 = std::min(bx0, v[k]); bx1 = std::max(bx1, v[k]);
                by0 = std::min(by0, v[k + 1]); by1 = std::max(by1, v[k + 1]);
            }
        }
    }
    return lef_error(L, ("PIN " + pname + " without END").c_str());
}

ep_status lef_macro(Lexer &L, ep_lefdef &d)
{
    std::string name;
    if (!L.next(name)) return lef_error(L, "MACRO without a name");
    Macro m;
    std::string t;
    while (L.next(t)) {
        if (t == "END") {
            std::string u;
            if (!L.next(u)) break;
            if (u == name) {
                d.macros[name] = std::move(m);
                return EP_OK;
            }
            continue;  // stray END (e.g. of a block this loop skipped)
        }
        if (t == "SIZE") {
            std::string a, by, b, semi;
            if (!L.next(a) || !L.next(by) || !L.next(b) || by != "BY" || !to_double(a, m.w) ||
                !to_double(b, m.h))
                return lef_error(L, ("malformed SIZE in MACRO " + name).c_str());
            skip_statement(L);
        } else if (t == "ORIGIN") {
            std::string a, b;
            if (!L.next(a) || !L.next(b) || !to_double(a, m.ox) || !to_double(b, m.oy))
                return lef_error(L, ("malformed ORIGIN in MACRO " + name).c_str());
            skip_statement(L);
        } else if (t == "PIN") {
            std::string pname;
            if (!L.next(pname)) return lef_error(L, "PIN without a name");
            ep_status st = lef_pin(L, pname, m);
            if (st != EP_OK) return st;
        } else if (t == "OBS") {
            if (!skip_to(L, "END")) return lef_error(L, "OBS without END");
        } else {
            if (!skip_statement(L)) break;
        }
    }
    return lef_error(L, ("MACRO " + name + " without END").c_str());
}

ep_status read_lef(const std::string &path, ep_lefdef &d)
{
    Lexer L;
    L.path = path;
    if (!read_file(path, L.text)) return fail(EP_EIO, "LEF %s: cannot read the file", path.c_str());
    std::string t;
    while (L.next(t)) {
        if (t == "MACRO") {
            ep_status st = lef_macro(L, d);
            if (st != EP_OK) return st;
        } else if (t == "LAYER" || t == "VIA" || t == "VIARULE" || t == "SITE" || t == "NONDEFAULTRULE" ||
                   t == "ARRAY") {
            std::string name;
            if (!L.next(name) || !skip_to(L, "END", name))
                return lef_error(L, (t + " block without END").c_str());
        } else if (t == "UNITS" || t == "PROPERTYDEFINITIONS" || t == "SPACING" || t == "BEGINEXT") {
            const std::string end = t == "BEGINEXT" ? "ENDEXT" : "END";
            if (!(end == "ENDEXT" ? skip_to(L, end) : skip_to(L, end, t)))
                return lef_error(L, (t + " block without END").c_str());
        } else if (t == "END") {
            std::string u;
            if (L.next(u) && u == "LIBRARY") break;
        } else {
            if (!skip_statement(L)) break;
        }
    }
    return EP_OK;
}

// ---------------------------------------------------------------- DEF

ep_status def_error(const Lexer &L, const char *what)
{
    return fail(EP_EIO, "DEF %s:%d: %s", L.path.c_str(), L.line, what);
}

// "( x y )" -> x, y (a '*' keeps the previous value).
bool def_point(Lexer &L, int64_t &x, int64_t &y)
{
    std::string a, b, c;
    if (!L.next(a) || a != "(" || !L.next(b) || !L.next(c)) return false;
    if (b != "*" && !to_int64(b, x)) return false;
    if (c != "*" && !to_int64(c, y)) return false;
    std::string e;
    while (L.next(e) && e != ")") {  // ( x y ext ) in routing; not used here
    }
    return true;
}

int status_of(const std::string &t)
{
    if (t == "PLACED") return kStatPlaced;
    if (t == "FIXED") return kStatFixed;
    if (t == "COVER") return kStatCover;
    if (t == "UNPLACED") return kStatUnplaced;
    return -1;
}

ep_status def_components(Lexer &L, ep_lefdef &d)
{
    std::string t;
    skip_statement(L);  // the count
    while (L.next(t)) {
        if (t == "END") {
            std::string u;
            L.next(u);
            return EP_OK;
        }
        if (t != "-") return def_error(L, ("unexpected '" + t + "' in COMPONENTS").c_str());
        Comp c;
        if (!L.next(c.name) || !L.next(c.macro)) return def_error(L, "component without name and macro");
        while (L.next(t) && t != ";") {
            if (t != "+") continue;
            std::string k;
            if (!L.next(k)) break;
            const int st = status_of(k);
            if (st < 0) continue;
            c.status = st;
            if (st == kStatUnplaced) {  // UNPLACED may carry a location
                const size_t save = L.at;
                const int line = L.line;
                std::string u;
                if (L.next(u) && u == "(") {
                    L.at = save;
                    L.line = line;
                } else {
                    L.at = save;
                    L.line = line;
                    continue;
                }
            }
            if (!def_point(L, c.x, c.y)) return def_error(L, ("bad location of " + c.name).c_str());
            c.has_loc = true;
            std::string o;
            if (!L.next(o)) break;
            if (o == ";") break;
            double qx, qy;
            if (!orient_point(o, 1, 1, 0, 0, qx, qy)) return def_error(L, ("bad orientation of " + c.name).c_str());
            c.orient = o;
        }
        d.comps.push_back(std::move(c));
    }
    return def_error(L, "COMPONENTS without END");
}

ep_status def_pins(Lexer &L, ep_lefdef &d)
{
    std::string t;
    skip_statement(L);
    while (L.next(t)) {
        if (t == "END") {
            std::string u;
            L.next(u);
            return EP_OK;
        }
        if (t != "-") return def_error(L, ("unexpected '" + t + "' in PINS").c_str());
        IoPin p;
        if (!L.next(p.name)) return def_error(L, "pin without a name");
        while (L.next(t) && t != ";") {
            if (t != "+") continue;
            std::string k;
            if (!L.next(k)) break;
            if (k == "NET") L.next(p.net);
            else if (k == "DIRECTION") L.next(p.dir);
            else if (k == "LAYER" && !p.has_box) {
                L.next(p.layer);
                std::string u;
                const size_t save = L.at;
                const int line = L.line;
                if (L.next(u) && u == "MASK") L.next(u);  // + LAYER l MASK n ( .. ) ( .. )
                else {
                    L.at = save;
                    L.line = line;
                }
                if (!def_point(L, p.box[0], p.box[1]) || !def_point(L, p.box[2], p.box[3]))
                    return def_error(L, ("bad LAYER box of pin " + p.name).c_str());
                p.has_box = true;
            } else {
                const int st = status_of(k);
                if (st < 0 || st == kStatUnplaced || p.has_loc) continue;
                p.status = st;
                if (!def_point(L, p.x, p.y)) return def_error(L, ("bad location of pin " + p.name).c_str());
                p.has_loc = true;
                std::string o;
                if (!L.next(o) || o == ";") break;
                p.orient = o;
            }
        }
        d.pins.push_back(std::move(p));
    }
    return def_error(L, "PINS without END");
}

ep_status def_nets(Lexer &L, ep_lefdef &d)
{
    std::unordered_map<std::string, int> comp_of, pin_of;
    for (size_t i = 0; i < d.comps.size(); ++i) comp_of.emplace(d.comps[i].name, (int)i);
    for (size_t i = 0; i < d.pins.size(); ++i) pin_of.emplace(d.pins[i].name, (int)i);
    std::string t;
    skip_statement(L);
    while (L.next(t)) {
        if (t == "END") {
            std::string u;
            L.next(u);
            return EP_OK;
        }
        if (t != "-") return def_error(L, ("unexpected '" + t + "' in NETS").c_str());
        DNet n;
        if (!L.next(n.name)) return def_error(L, "net without a name");
        bool conns = true;  // connections come before the first '+'
        while (L.next(t) && t != ";") {
            if (t == "+") conns = false;
            if (!conns || t != "(") continue;
            std::string a, b;
            if (!L.next(a) || !L.next(b)) return def_error(L, ("bad connection in net " + n.name).c_str());
            std::string e;
            while (L.next(e) && e != ")") {  // ( comp pin + SYNTHESIZED )
            }
            Conn c;
            if (a == "PIN") {
                auto it = pin_of.find(b);
                if (it == pin_of.end())
                    return fail(EP_EINVAL, "DEF %s:%d: net %s names unknown IO pin %s", L.path.c_str(), L.line,
                                n.name.c_str(), b.c_str());
                c.io = 1;
                c.idx = it->second;
            } else if (a == "*") {
                continue;  // ( * pin ): every component's pin of that name (global nets); no placement pin
            } else {
                auto it = comp_of.find(a);
                if (it == comp_of.end())
                    return fail(EP_EINVAL, "DEF %s:%d: net %s names unknown component %s", L.path.c_str(), L.line,
                                n.name.c_str(), a.c_str());
                c.idx = it->second;
                c.pin = b;
            }
            n.conn.push_back(std::move(c));
        }
        d.nets.push_back(std::move(n));
    }
    return def_error(L, "NETS without END");
}

ep_status read_def(const std::string &path, ep_lefdef &d)
{
    Lexer L;
    L.path = path;
    if (!read_file(path, L.text)) return fail(EP_EIO, "DEF %s: cannot read the file", path.c_str());
    std::string t;
    while (L.next(t)) {
        ep_status st = EP_OK;
        if (t == "DESIGN") {
            if (!L.next(d.name)) return def_error(L, "DESIGN without a name");
            skip_statement(L);
        } else if (t == "UNITS") {
            std::string a, b, c;
            int64_t u = 0;
            if (!L.next(a) || !L.next(b) || !L.next(c) || a != "DISTANCE" || b != "MICRONS" || !to_int64(c, u) ||
                u <= 0)
                return def_error(L, "malformed UNITS");
            d.units = (int32_t)u;
            skip_statement(L);
        } else if (t == "DIEAREA") {
            int64_t x0 = INT64_MAX, y0 = INT64_MAX, x1 = INT64_MIN, y1 = INT64_MIN, px = 0, py = 0;
            std::string u;
            int n = 0;
            for (;;) {
                const size_t save = L.at;
                const int line = L.line;
                if (!L.next(u)) return def_error(L, "DIEAREA without ';'");
                if (u == ";") break;
                L.at = save;
                L.line = line;
                if (!def_point(L, px, py)) return def_error(L, "malformed DIEAREA");
                x0 = std::min(x0, px); y0 = std::min(y0, py);
                x1 = std::max(x1, px); y1 = std::max(y1, py);
                ++n;
            }
            if (n < 2 || !(x1 > x0) || !(y1 > y0)) return def_error(L, "empty DIEAREA");
            d.die[0] = x0; d.die[1] = y0; d.die[2] = x1; d.die[3] = y1;
            d.has_die = true;
        } else if (t == "COMPONENTS") {
            st = def_components(L, d);
        } else if (t == "PINS") {
            st = def_pins(L, d);
        } else if (t == "NETS") {
            st = def_nets(L, d);
        } else if (t == "PROPERTYDEFINITIONS" || t == "VIAS" || t == "NONDEFAULTRULES" || t == "REGIONS" ||
                   t == "GROUPS" || t == "BLOCKAGES" || t == "SPECIALNETS" || t == "FILLS" || t == "SCANCHAINS" ||
                   t == "STYLES" || t == "SLOTS" || t == "PINPROPERTIES" || t == "COMPONENTMASKSHIFT" ||
                   t == "HISTORY" || t == "BEGINEXT") {
            const bool ext = t == "BEGINEXT";
            if (t == "COMPONENTMASKSHIFT" || t == "HISTORY") {
                skip_statement(L);
            } else if (!(ext ? skip_to(L, "ENDEXT") : skip_to(L, "END", t))) {
                return def_error(L, (t + " without END").c_str());
            }
        } else if (t == "END") {
            std::string u;
            if (L.next(u) && u == "DESIGN") break;
        } else {
            if (!skip_statement(L)) break;
        }
        if (st != EP_OK) return st;
    }
    if (!d.has_die) return def_error(L, "no DIEAREA");
    return EP_OK;
}

// ---------------------------------------------------------------- netlist view

ep_status build_view(ep_lefdef &d)
{
    const double u = (double)d.units;
    const double cx = 0.5 * (d.die[0] + d.die[2]) / u, cy = 0.5 * (d.die[1] + d.die[3]) / u;
    std::vector<int> mov, fix;
    for (size_t i = 0; i < d.comps.size(); ++i) {
        if (!d.macros.count(d.comps[i].macro))
            return fail(EP_EINVAL, "DEF: component %s is of unknown macro %s", d.comps[i].name.c_str(),
                        d.comps[i].macro.c_str());
        (is_fixed(d.comps[i].status) ? fix : mov).push_back((int)i);
    }
    d.n_mov = (int32_t)mov.size();
    d.n_fix = (int32_t)(fix.size() + d.pins.size());
    const size_t n_all = mov.size() + fix.size() + d.pins.size();
    d.w.assign(n_all, 0.f); d.h.assign(n_all, 0.f); d.x.assign(n_all, 0.f); d.y.assign(n_all, 0.f);
    d.cell_of_comp.assign(d.comps.size(), -1);
    d.cell_of_pin.assign(d.pins.size(), -1);
    d.movable_comps = mov;
    int32_t k = 0;
    auto put_comp = [&](int i) {
        const Comp &c = d.comps[i];
        const Macro &m = d.macros.at(c.macro);
        const bool q = quarter_turn(c.orient);
        d.w[k] = (float)(q ? m.h : m.w);
        d.h[k] = (float)(q ? m.w : m.h);
        if (c.has_loc) {
            d.x[k] = (float)(c.x / u);
            d.y[k] = (float)(c.y / u);
        } else {  // no location: the die centre
            d.x[k] = (float)(cx - 0.5 * d.w[k]);
            d.y[k] = (float)(cy - 0.5 * d.h[k]);
        }
        d.cell_of_comp[i] = k++;
    };
    for (int i : mov) put_comp(i);
    for (int i : fix) put_comp(i);
    const float tiny = (float)(1.0 / u);  // a pin without a LAYER box: one database unit square
    for (size_t i = 0; i < d.pins.size(); ++i) {
        const IoPin &p = d.pins[i];
        double bx0 = 0, by0 = 0, bx1 = 0, by1 = 0;
        if (p.has_box) {
            double qx0, qy0, qx1, qy1;
            // box corners relative to the pin location, under its orientation (about the origin)
            orient_point(p.orient, 0, 0, (double)p.box[0], (double)p.box[1], qx0, qy0);
            orient_point(p.orient, 0, 0, (double)p.box[2], (double)p.box[3], qx1, qy1);
            bx0 = std::min(qx0, qx1); bx1 = std::max(qx0, qx1);
            by0 = std::min(qy0, qy1); by1 = std::max(qy0, qy1);
        }
        const double ox = p.has_loc ? (double)p.x : 0.0, oy = p.has_loc ? (double)p.y : 0.0;
        d.w[k] = std::max(tiny, (float)((bx1 - bx0) / u));
        d.h[k] = std::max(tiny, (float)((by1 - by0) / u));
        d.x[k] = (float)((ox + bx0) / u);
        d.y[k] = (float)((oy + by0) / u);
        d.cell_of_pin[i] = k++;
    }
    // nets of two or more pins
    d.net_start.assign(1, 0);
    d.pin_cell.clear(); d.pin_dx.clear(); d.pin_dy.clear();
    for (const DNet &n : d.nets) {
        if (n.conn.size() < 2) continue;
        for (const Conn &c : n.conn) {
            int32_t cell;
            double dx, dy;
            if (c.io) {
                cell = d.cell_of_pin[c.idx];
                dx = 0.5 * d.w[cell];
                dy = 0.5 * d.h[cell];
            } else {
                const Comp &cm = d.comps[c.idx];
                const Macro &m = d.macros.at(cm.macro);
                cell = d.cell_of_comp[c.idx];
                auto it = m.pins.find(c.pin);
                const double px = it == m.pins.end() ? 0.5 * m.w : it->second.first + m.ox;
                const double py = it == m.pins.end() ? 0.5 * m.h : it->second.second + m.oy;
                orient_point(cm.orient, m.w, m.h, px, py, dx, dy);
            }
            d.pin_cell.push_back(cell);
            d.pin_dx.push_back((float)dx);
            d.pin_dy.push_back((float)dy);
        }
        d.net_start.push_back((int32_t)d.pin_cell.size());
    }
    return EP_OK;
}

ep_status finish(ep_lefdef *d, ep_status st, ep_lefdef **out)
{
    if (st == EP_OK) st = build_view(*d);
    if (st != EP_OK) {
        delete d;
        return st;
    }
    *out = d;
    return EP_OK;
}

std::string join(const std::string &dir, const std::string &name)
{
    if (dir.empty() || dir.back() == '/') return dir + name;
    return dir + "/" + name;
}

}  // namespace

extern "C" ep_status ep_lefdef_read(const char *const *lef_paths, int32_t n_lef, const char *def_path,
                                    ep_lefdef **out)
{
    if (!out || !def_path || n_lef < 0 || (n_lef > 0 && !lef_paths))
        return fail(EP_EINVAL, "ep_lefdef_read: null argument");
    *out = nullptr;
    ep_lefdef *d = new (std::nothrow) ep_lefdef();
    if (!d) return fail(EP_ENOMEM, "ep_lefdef_read: host allocation failed");
    ep_status st = EP_OK;
    for (int i = 0; i < n_lef && st == EP_OK; ++i) {
        if (!lef_paths[i]) st = fail(EP_EINVAL, "ep_lefdef_read: null LEF path %d", i);
        else st = read_lef(lef_paths[i], *d);
    }
    if (st == EP_OK) st = read_def(def_path, *d);
    return finish(d, st, out);
}

extern "C" ep_status ep_lefdef_read_dir(const char *dir, ep_lefdef **out)
{
    if (!out || !dir) return fail(EP_EINVAL, "ep_lefdef_read_dir: null argument");
    *out = nullptr;
    std::vector<std::string> lefs, defs;
    DIR *dh = opendir(dir);
    if (!dh) return fail(EP_EIO, "ep_lefdef_read_dir: cannot open %s", dir);
    while (dirent *e = readdir(dh)) {
        const std::string n = e->d_name;
        auto ends = [&](const char *x) {
            const size_t k = std::strlen(x);
            return n.size() > k && n.compare(n.size() - k, k, x) == 0;
        };
        if (ends(".lef")) lefs.push_back(n);
        else if (ends(".def")) defs.push_back(n);
    }
    closedir(dh);
    std::sort(lefs.begin(), lefs.end());
    if (lefs.empty()) return fail(EP_EIO, "ep_lefdef_read_dir: no .lef files in %s", dir);
    std::string def;
    if (std::find(defs.begin(), defs.end(), "floorplan.def") != defs.end()) def = "floorplan.def";
    else if (defs.size() == 1) def = defs[0];
    else return fail(EP_EIO, "ep_lefdef_read_dir: %s has no floorplan.def (and %zu other .def files)", dir, defs.size());
    std::vector<std::string> full;
    for (const std::string &l : lefs) full.push_back(join(dir, l));
    std::vector<const char *> ptr;
    for (const std::string &f : full) ptr.push_back(f.c_str());
    const std::string dpath = join(dir, def);
    return ep_lefdef_read(ptr.data(), (int32_t)ptr.size(), dpath.c_str(), out);
}

extern "C" ep_status ep_lefdef_free(ep_lefdef *d)
{
    if (!d) return fail(EP_EINVAL, "ep_lefdef_free: null design");
    delete d;
    return EP_OK;
}

extern "C" ep_status ep_lefdef_netlist(const ep_lefdef *d, ep_netlist *nl)
{
    if (!d || !nl) return fail(EP_EINVAL, "ep_lefdef_netlist: null argument");
    const double u = (double)d->units;
    nl->n_mov = d->n_mov;
    nl->n_fix = d->n_fix;
    nl->w = d->w.data(); nl->h = d->h.data();
    nl->x = d->x.data(); nl->y = d->y.data();
    nl->n_nets = (int32_t)d->net_start.size() - 1;
    nl->n_pins = (int32_t)d->pin_cell.size();
    nl->net_start = d->net_start.data();
    nl->pin_cell = d->pin_cell.data();
    nl->pin_dx = d->pin_dx.data(); nl->pin_dy = d->pin_dy.data();
    nl->xl = (float)(d->die[0] / u); nl->yl = (float)(d->die[1] / u);
    nl->xh = (float)(d->die[2] / u); nl->yh = (float)(d->die[3] / u);
    return EP_OK;
}

extern "C" ep_status ep_lefdef_get_info(const ep_lefdef *d, ep_lefdef_info *info, char *design_name)
{
    if (!d || !info) return fail(EP_EINVAL, "ep_lefdef_get_info: null argument");
    const double u = (double)d->units;
    info->units = d->units;
    info->n_macros = (int32_t)d->macros.size();
    info->n_components = (int32_t)d->comps.size();
    info->n_io_pins = (int32_t)d->pins.size();
    info->n_nets = (int32_t)d->nets.size();
    info->n_mov = d->n_mov;
    info->n_fix = d->n_fix;
    info->view_nets = (int32_t)d->net_start.size() - 1;
    info->view_pins = (int32_t)d->pin_cell.size();
    for (int k = 0; k < 4; ++k) info->die[k] = (double)d->die[k] / u;
    if (design_name) {
        std::strncpy(design_name, d->name.c_str(), 255);
        design_name[255] = '\0';
    }
    return EP_OK;
}

extern "C" ep_status ep_lefdef_write_def(const ep_lefdef *d, const float *x, const float *y,
                                         const char *output_dir)
{
    if (!d || !output_dir || (d->n_mov > 0 && (!x || !y)))
        return fail(EP_EINVAL, "ep_lefdef_write_def: null argument");
    for (int32_t k = 0; k < d->n_mov; ++k)
        if (!std::isfinite(x[k]) || !std::isfinite(y[k]))
            return fail(EP_ENAN, "ep_lefdef_write_def: non-finite position of cell %d", k);
    const std::string path = join(output_dir, d->name + ".def");
    std::ofstream out(path);
    if (!out) return fail(EP_EIO, "ep_lefdef_write_def: cannot write %s", path.c_str());
    out.imbue(std::locale::classic());
    const double u = (double)d->units;
    std::vector<int64_t> px(d->comps.size()), py(d->comps.size());
    for (size_t i = 0; i < d->comps.size(); ++i) {
        px[i] = d->comps[i].x;
        py[i] = d->comps[i].y;
    }
    for (int32_t k = 0; k < d->n_mov; ++k) {
        const int i = d->movable_comps[k];
        px[i] = (int64_t)std::llround((double)x[k] * u);
        py[i] = (int64_t)std::llround((double)y[k] * u);
    }
    out << "# Design after global placement\n# Produced by aieplace_amd\n";
    out << "VERSION 5.8 ;\nDIVIDERCHAR \"/\" ;\nBUSBITCHARS \"[]\" ;\n";
    out << "DESIGN " << d->name << " ;\nUNITS DISTANCE MICRONS " << d->units << " ;\n\n";
    out << "DIEAREA ( " << d->die[0] << " " << d->die[1] << " ) ( " << d->die[2] << " " << d->die[3] << " ) ;\n\n";
    out << "COMPONENTS " << d->comps.size() << " ;\n";
    for (size_t i = 0; i < d->comps.size(); ++i) {
        const Comp &c = d->comps[i];
        // placed by this run: movable components are PLACED
        const int st = is_fixed(c.status) ? c.status : kStatPlaced;
        out << "    - " << c.name << " " << c.macro << "\n      + " << status_word(st) << " ( " << px[i] << " "
            << py[i] << " ) " << c.orient << " ;\n";
    }
    out << "END COMPONENTS\n\n";
    out << "PINS " << d->pins.size() << " ;\n";
    for (const IoPin &p : d->pins) {
        out << "    - " << p.name << " + NET " << (p.net.empty() ? p.name : p.net);
        if (!p.dir.empty()) out << "\n      + DIRECTION " << p.dir;
        if (p.has_loc) out << "\n      + " << status_word(p.status) << " ( " << p.x << " " << p.y << " ) " << p.orient;
        if (p.has_box)
            out << "\n      + LAYER " << p.layer << " ( " << p.box[0] << " " << p.box[1] << " ) ( " << p.box[2] << " "
                << p.box[3] << " )";
        out << " ;\n";
    }
    out << "END PINS\n\n";
    out << "NETS " << d->nets.size() << " ;\n";
    for (const DNet &n : d->nets) {
        out << "    - " << n.name;
        int count = 0;
        for (const Conn &c : n.conn) {
            if (c.io) out << " ( PIN " << d->pins[c.idx].name << " )";
            else out << " ( " << d->comps[c.idx].name << " " << c.pin << " )";
            if (++count == 4 && &c != &n.conn.back()) {  // four connections per line
                out << "\n     ";
                count = 0;
            }
        }
        out << " ;\n";
    }
    out << "END NETS\n\nEND DESIGN\n";
    out.flush();
    if (!out) return fail(EP_EIO, "ep_lefdef_write_def: write to %s failed", path.c_str());
    return EP_OK;
}
