// JSON config reader with the reference's keys, types and defaults
// (source/config.d:112-313).  "dt" must be a JSON float; "saveIter"/"plotIter"
// integers; the boundary list key is spelled "boudaryConditions"; unknown BC type
// strings leave the default type FullState.  Extension keys live under the
// "extensions" object, which reference configs do not have.
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>

#include "host.hpp"

namespace ebb {

namespace {

struct JVal {
    enum Kind { Null, True, False, Int, Float, Str, Arr, Obj } kind = Null;
    long long i = 0;
    double f = 0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    const JVal* get(const std::string& k) const
    {
        if (kind != Obj) return nullptr;
        for (auto& kv : o) if (kv.first == k) return &kv.second;
        return nullptr;
    }
    bool isNum() const { return kind == Int || kind == Float; }
    double num() const { return kind == Int ? double(i) : f; }
};

struct Parser {
    const std::string& t;
    size_t p = 0;
    int depth = 0;                       // nesting of the value being parsed
    static constexpr int MAX_DEPTH = 64; // configs nest 3 deep; deeper input is rejected, not recursed into
    std::string err;
    explicit Parser(const std::string& s) : t(s) {}
    void ws() { while (p < t.size() && std::isspace((unsigned char)t[p])) p++; }
    bool fail(const std::string& m) { if (err.empty()) err = m + " at offset " + std::to_string(p); return false; }
    bool str(std::string& out)
    {
        if (t[p] != '"') return fail("expected string");
        p++;
        out.clear();
        while (p < t.size() && t[p] != '"') {
            if (t[p] == '\\') {
                p++;
                if (p >= t.size()) return fail("bad escape");
                char c = t[p];
                if (c == 'n') out.push_back('\n');
                else if (c == 't') out.push_back('\t');
                else if (c == 'u') { p += 4; out.push_back('?'); }
                else out.push_back(c);
                p++;
            } else {
                out.push_back(t[p++]);
            }
        }
        if (p >= t.size()) return fail("unterminated string");
        p++;
        return true;
    }
    bool val(JVal& v)
    {
        if (depth >= MAX_DEPTH) return fail("nesting too deep");
        struct Level {
            int& d;
            explicit Level(int& x) : d(x) { ++d; }
            ~Level() { --d; }
        } level(depth);
        ws();
        if (p >= t.size()) return fail("unexpected end");
        char c = t[p];
        if (c == '{') {
            v.kind = JVal::Obj;
            p++;
            ws();
            if (t[p] == '}') { p++; return true; }
            while (true) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (t[p] != ':') return fail("expected ':'");
                p++;
                JVal x;
                if (!val(x)) return false;
                v.o.emplace_back(k, std::move(x));
                ws();
                if (t[p] == ',') { p++; continue; }
                if (t[p] == '}') { p++; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JVal::Arr;
            p++;
            ws();
            if (t[p] == ']') { p++; return true; }
            while (true) {
                JVal x;
                if (!val(x)) return false;
                v.a.push_back(std::move(x));
                ws();
                if (t[p] == ',') { p++; continue; }
                if (t[p] == ']') { p++; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.kind = JVal::Str; return str(v.s); }
        if (t.compare(p, 4, "true") == 0) { v.kind = JVal::True; p += 4; return true; }
        if (t.compare(p, 5, "false") == 0) { v.kind = JVal::False; p += 5; return true; }
        if (t.compare(p, 4, "null") == 0) { v.kind = JVal::Null; p += 4; return true; }
        size_t s = p;
        bool isFloat = false;
        if (t[p] == '-' || t[p] == '+') p++;
        while (p < t.size() && (std::isdigit((unsigned char)t[p]) || t[p] == '.' || t[p] == 'e' || t[p] == 'E' || t[p] == '-' || t[p] == '+')) {
            if (t[p] == '.' || t[p] == 'e' || t[p] == 'E') isFloat = true;
            p++;
        }
        if (s == p) return fail("unexpected character");
        std::string num = t.substr(s, p - s);
        if (isFloat) { v.kind = JVal::Float; v.f = std::strtod(num.c_str(), nullptr); }
        else { v.kind = JVal::Int; v.i = std::strtoll(num.c_str(), nullptr, 10); }
        return true;
    }
};

int bcType(const std::string& s, bool& known)
{
    known = true;
    if (s == "fullState") return FullState;
    if (s == "inviscidWall") return InviscidWall;
    if (s == "viscidWall") return ViscousWall;
    if (s == "constP") return ConstPressure;
    if (s == "symmetry") return Symmetry;
    if (s == "tpInflow") return TempPresInflow;
    if (s == "dirichlet") return Dirichlet;
    known = false;
    return FullState;
}

bool isTrue(const JVal* v) { return v && v->kind == JVal::True; }

}  // namespace

bool parseConfig(const std::string& json, Config& c, std::string& err)
{
    Parser P(json);
    JVal root;
    if (!P.val(root)) { err = "JSON parse error: " + P.err; return false; }
    if (root.kind != JVal::Obj) { err = "config must be a JSON object"; return false; }
    c = Config();
    auto need = [&](const char* k) -> const JVal* {
        const JVal* v = root.get(k);
        if (!v) err = std::string("missing required key \"") + k + "\"";
        return v;
    };
    const JVal* v;
    if (!(v = need("mesh")) || v->kind != JVal::Str) { if (err.empty()) err = "\"mesh\" must be a string"; return false; }
    c.meshFile = v->s;
    if (!(v = need("flux")) || v->kind != JVal::Str) { if (err.empty()) err = "\"flux\" must be a string"; return false; }
    c.flux = v->s;
    if (c.flux != "roeFlux" && c.flux != "rusanovFlux") { err = "unknown flux " + c.flux; return false; }
    if (!(v = need("dt")) || v->kind != JVal::Float) { if (err.empty()) err = "\"dt\" must be a JSON float"; return false; }
    c.dt = v->f;
    if (!(v = need("tEnd")) || !v->isNum()) { if (err.empty()) err = "\"tEnd\" must be a number"; return false; }
    c.tEnd = v->num();
    if (!(v = need("saveIter")) || v->kind != JVal::Int) { if (err.empty()) err = "\"saveIter\" must be an integer"; return false; }
    c.saveIter = v->i;
    if (!(v = need("plotIter")) || v->kind != JVal::Int) { if (err.empty()) err = "\"plotIter\" must be an integer"; return false; }
    c.plotIter = v->i;
    if (!(v = need("forceBoundary"))) return false;
    if (v->kind == JVal::Arr) { for (auto& s : v->a) c.forceBoundary.push_back(s.s); }
    else if (v->kind == JVal::Str) c.forceBoundary.push_back(v->s);
    else { err = "\"forceBoundary\" must be a string or array"; return false; }
    if ((v = root.get("CFL")) && v->isNum()) c.CFL = v->num();
    else c.warnings.push_back("CFL not provided, setting to 0.5");
    if ((v = root.get("dynamicDt"))) c.dynamicDt = isTrue(v);
    if ((v = root.get("integrator")) && v->kind == JVal::Str) c.integrator = v->s;
    if (c.integrator != "Euler" && c.integrator != "RK2" && c.integrator != "RK2_TVD" && c.integrator != "RK4") {
        err = "unknown integrator " + c.integrator;
        return false;
    }
    if ((v = root.get("limited"))) c.limited = isTrue(v);
    if ((v = root.get("localTimestep"))) c.localTimestep = isTrue(v);
    if ((v = root.get("lpThresh")) && v->isNum()) c.lpThresh = v->num();
    if ((v = root.get("order")) && v->kind == JVal::Int) {
        c.order = v->i;
        if (c.order != 1 && c.order != 2) c.order = 2;
    }
    if ((v = root.get("multistageLimiting"))) c.multistageLimiting = isTrue(v);
    if ((v = root.get("cflAdjust"))) c.cflAdjust = isTrue(v);
    if ((v = root.get("viscosity"))) c.viscosity = isTrue(v);
    if ((v = root.get("physicalConfig")) && v->kind == JVal::Obj) {
        const JVal *a = v->get("Pr"), *b = v->get("R"), *g = v->get("gamma"), *r = v->get("Re"), *l = v->get("L");
        if (a && b && g && r && l && a->isNum() && b->isNum() && g->isNum() && r->isNum() && l->isNum()) {
            c.Pr = a->num(); c.R = b->num(); c.gamma = g->num(); c.Re = r->num(); c.L = l->num();
        }
    }
    if (!(v = need("initialConditions")) || v->kind != JVal::Arr || v->a.size() < 4) {
        if (err.empty()) err = "\"initialConditions\" must have 4 numbers";
        return false;
    }
    for (int i = 0; i < 4; i++) c.ic[i] = v->a[i].num();
    if (!(v = need("boudaryConditions")) || v->kind != JVal::Arr) { if (err.empty()) err = "\"boudaryConditions\" must be an array"; return false; }
    for (auto& b : v->a) {
        BCConfig bc;
        const JVal* t = b.get("tag");
        const JVal* ty = b.get("type");
        const JVal* q = b.get("q");
        if (!t || !ty || !q || q->kind != JVal::Arr) { err = "boundary condition needs tag, type and q"; return false; }
        bc.tag = t->s;
        bool known;
        bc.type = bcType(ty->s, known);
        if (!known) c.warnings.push_back("unknown boundary type \"" + ty->s + "\" treated as fullState");
        for (auto& x : q->a) bc.data.push_back(x.num());
        c.bcs.push_back(bc);
    }
    if ((v = root.get("extensions")) && v->kind == JVal::Obj) {
        const JVal* e;
        if ((e = v->get("mms"))) c.mms = isTrue(e);
        if ((e = v->get("maxSteps")) && e->kind == JVal::Int) c.maxSteps = e->i;
        if ((e = v->get("rmsDrop")) && e->isNum()) {
            if (!(e->num() >= 0 && e->num() < 1)) { err = "rmsDrop must be in [0, 1)"; return false; }
            c.rmsDrop = e->num();
        }
        if ((e = v->get("cellOrder")) && e->kind == JVal::Str) {
            if (e->s == "reference") c.renumber = false;
            else if (e->s == "tiles") c.renumber = true;
            else { err = "unknown cellOrder " + e->s; return false; }
        }
        if ((e = v->get("stagePath")) && e->kind == JVal::Str) {
            if (e->s == "split") c.stagePath = 0;
            else if (e->s == "faceTiles") c.stagePath = 1;
            else if (e->s == "fused") c.stagePath = 2;
            else if (e->s == "gradTiles") c.stagePath = 3;
            else { err = "unknown stagePath " + e->s; return false; }
        }
        if ((e = v->get("tileCells")) && e->kind == JVal::Int) {
            if (e->i != -1 && (e->i < 1 || e->i > 1024)) { err = "tileCells must be -1 (automatic) or in 1..1024"; return false; }
            c.tileCells = e->i;
        }
        if ((e = v->get("kernelOrder")) && e->kind == JVal::Str) {
            if (e->s == "forward") c.kernelOrder = 0;
            else if (e->s == "tilesBackward") c.kernelOrder = 1;
            else if (e->s == "gradientBackward") c.kernelOrder = 2;
            else if (e->s == "bothBackward") c.kernelOrder = 3;
            else { err = "unknown kernelOrder " + e->s; return false; }
        }
        if ((e = v->get("tileOrder")) && e->kind == JVal::Str) {
            if (e->s == "grow") c.tileOrder = 0;
            else if (e->s == "morton") c.tileOrder = 1;
            else if (e->s == "blocks") c.tileOrder = 2;
            else { err = "unknown tileOrder " + e->s; return false; }
        }
        if ((e = v->get("limiter")) && e->kind == JVal::Str) {
            if (e->s == "scalar") c.limiter = 0;
            else if (e->s == "venkatakrishnan") c.limiter = 1;
            else { err = "unknown limiter " + e->s; return false; }
        }
        if ((e = v->get("venkatK")) && e->isNum()) {
            if (!(e->num() >= 0)) { err = "venkatK must be >= 0"; return false; }
            c.venkatK = e->num();
        }
        if ((e = v->get("precision")) && e->kind == JVal::Str) {
            if (e->s == "mixed") c.precision = 1;
            else if (e->s == "fp64") c.precision = 0;
            else { err = "unknown precision " + e->s; return false; }
        }
    }
    return true;
}

}  // namespace ebb
