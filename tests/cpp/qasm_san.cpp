// Runs the OpenQASM reader (qi_qasm_parse) on every file named on the command
// line: the size query, then the parse into buffers of exactly that size.
// Built with the address and undefined-behaviour sanitizers by
// tests/test_qasm.py, on valid programs and on truncated and corrupted ones:
// malformed input must come back as an error, never as a memory fault.
// Prints one line per file: "ok <gates> <qubits> <params>" or "err <code> <line>".
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qiron.h"

int main(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
        FILE *f = fopen(argv[i], "rb");
        if (!f) { printf("missing\n"); continue; }
        std::string text;
        char buf[4096];
        size_t k;
        while ((k = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
        fclose(f);
        // no terminating zero: the reader must stay inside len
        std::vector<char> src(text.begin(), text.end());
        size_t ng = 0;
        uint32_t nq = 0, np = 0;
        qi_error err{};
        qi_status st = qi_qasm_parse(src.data(), src.size(), nullptr, 0, nullptr, nullptr, 0, &ng, &nq, &np, &err);
        if (st != QI_OK) { printf("err %d %llu\n", (int)st, (unsigned long long)err.a); continue; }
        std::vector<qi_gate> gates(ng);
        std::vector<int32_t> slots(ng);
        std::vector<double> vals(np, 0.25);
        st = qi_qasm_parse(src.data(), src.size(), vals.data(), vals.size(), gates.data(), slots.data(), ng, &ng, &nq,
                           &np, &err);
        if (st != QI_OK) { printf("err %d %llu\n", (int)st, (unsigned long long)err.a); continue; }
        printf("ok %zu %u %u\n", ng, nq, np);
    }
    return 0;
}
