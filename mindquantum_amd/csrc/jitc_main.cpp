// jitc_main.cpp - mqamd_jitc, the compile helper of the code generator (jit.cpp).
//
// hipRTC compiles one program at a time inside a process, so the library starts several of
// these helpers to compile the passes of a circuit in parallel.
// Usage: mqamd_jitc OUT1 SRC1 [OUT2 SRC2 ...]; on failure the compiler log goes to OUT.log.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "jit.h"

int main(int argc, char** argv) {
    if (argc < 3 || (argc - 1) % 2 != 0) {
        std::fprintf(stderr, "usage: mqamd_jitc OUT SRC [OUT SRC ...]\n");
        return 2;
    }
    int status = 0;
    for (int i = 1; i + 1 < argc; i += 2) {
        std::ifstream in(argv[i + 1]);
        std::stringstream ss;
        ss << in.rdbuf();
        std::string code, log;
        if (!in || !mqa::jit_compile(ss.str(), code, log)) {
            std::ofstream(std::string(argv[i]) + ".log") << (in ? log : std::string("cannot read ") + argv[i + 1]);
            status = 1;
            continue;
        }
        std::ofstream out(argv[i], std::ios::binary);
        out.write(code.data(), std::streamsize(code.size()));
        if (!out) status = 1;
    }
    return status;
}
