// Parses the TOML inputs named on the command line and a set of malformed
// snippets with the drop-in's parser (cahn-hilliard_amd/host/toml_lite.cpp),
// so that an AddressSanitizer / UndefinedBehaviorSanitizer build reports any
// invalid access.  Test infrastructure: built and run by tests/test_sanitizers.py.
#include <cstdio>

#include "toml_lite.h"

int main(int argc, char **argv) {
    int ok = 0;
    for(int i = 1; i < argc; i++) {
        chtoml::Table t;
        std::string err;
        if(!chtoml::parse_file(argv[i], t, err)) {
            std::fprintf(stderr, "%s: %s\n", argv[i], err.c_str());
            return 1;
        }
        ok += (int)t.size();
        if(chtoml::find(t, "mobility.type")) ok++;
    }
    const char *bad[] = {"a = ", "[x", "a = \"unterminated", "a = [1, 2", "a = { b = 1", "= 3", "a = 1e", "[[", "a.b.",
                         "a = 'x", "a = [1, [2, 3]]", "a = tru", "a = 0x", "a = \"\\u12\"", "a = {b = [1,}"};
    for(const char *s : bad) {
        chtoml::Table t;
        std::string err;
        chtoml::parse(s, t, err);
    }
    std::printf("parsed %d\n", ok);
    return 0;
}
