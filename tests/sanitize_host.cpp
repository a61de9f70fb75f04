// Host-only driver for tests/test_host_sanitizers.py: reads each mesh (gmsh 2.2 or .gri)
// and builds its geometry, or parses each .json config, with the library's own host
// sources compiled under AddressSanitizer and UndefinedBehaviorSanitizer.  Prints one
// line per input; files whose name contains "trunc" are expected to be rejected.
// Exit status: the number of inputs that were not handled as expected.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "../ebbcfd_amd/csrc/host.hpp"

using namespace ebb;

static std::string slurp(const char* p)
{
    std::ifstream f(p);
    std::stringstream s;
    s << f.rdbuf();
    return s.str();
}

int main(int argc, char** argv)
{
    int bad = 0;
    for (int a = 1; a < argc; a++) {
        const std::string p = argv[a];
        const bool expectFail = p.find("trunc") != std::string::npos;
        std::string err;
        bool ok;
        if (p.size() > 5 && p.substr(p.size() - 5) == ".json") {
            Config c;
            ok = parseConfig(slurp(argv[a]), c, err);
        } else {
            RawMesh raw;
            ok = p.find(".gri") != std::string::npos ? parseGri(p, raw, err) : parseGmsh(p, raw, err);
            BuiltMesh b;
            if (ok) ok = buildMesh(raw, b, err);
        }
        std::printf("%s %s %s\n", ok ? "ok" : "rejected", argv[a], err.c_str());
        if (ok == expectFail) bad++;
    }
    return bad;
}
