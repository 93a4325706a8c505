// Test harness: reads a PLY with the command line's reader and writes it back with its writer.
//   ply_roundtrip in.ply out.ply
#include <cstdio>

#include "ply_io.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    try {
        Mesh m;
        ply::read(argv[1], &m);
        std::printf("%zu %zu %zu %zu\n", m.nV(), m.nF(), m.nN(), m.nC());
        ply::write(argv[2], &m);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
