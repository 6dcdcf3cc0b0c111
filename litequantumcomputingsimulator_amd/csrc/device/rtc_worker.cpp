// lqcs_rtc_worker: compiles one specialized pass for the engine in a process
// of its own. hipRTC serialises the compiles of one process, so the engine
// compiles the distinct passes of a plan in several of these at once. Reads the
// generated kernel text on stdin, writes the gfx950 code object on stdout;
// exit status 1 with the compiler log on stderr when the compile fails. It
// makes no HIP runtime call and never touches a device.
#include "spec_compile.hpp"

#include <cstdio>
#include <iterator>
#include <iostream>
#include <string>
#include <vector>

namespace {
const char* const kSpecHeader =
#include "spec_pass_hpp.inc"
    ;
}  // namespace

int main() {
    std::ios::sync_with_stdio(false);
    const std::string body((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
    if (body.empty()) {
        std::fputs("lqcs_rtc_worker: no kernel text on stdin\n", stderr);
        return 1;
    }
    std::vector<char> code;
    std::string log;
    if (!lqcs_rtc::compile_pass(kSpecHeader, body, code, log)) {
        std::fputs(log.c_str(), stderr);
        return 1;
    }
    return std::fwrite(code.data(), 1, code.size(), stdout) == code.size() && std::fflush(stdout) == 0 ? 0 : 1;
}
