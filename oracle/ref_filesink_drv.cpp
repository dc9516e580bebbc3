// Driver for the reference's own FileSink (src/FileSink.cpp:30-107), built
// from the reference sources into oracle/_ref/ by `make -C oracle ref`.  It
// is test infrastructure: the sink tests write the same float arrays through
// it and through swspec_amd/sinks.py and compare the files byte for byte.
//
//   filesink_ref <Binary|Ascii> <float32 input file> <out file> <len bytes> [<len bytes> ...]
//
// Each length is one FileSink::write of a Buffer holding the next `len` bytes
// of the input (a Buffer allocated with room to spare, length set by
// setLength, as TaskCoreIPP fills them); a length of 0 writes an empty Buffer.
#include "Buffer.h"
#include "FileSink.h"
#include "Settings.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s Binary|Ascii in.f32 out len...\n", argv[0]);
        return 2;
    }
    swspect_settings_t st;
    st.sinkformat = std::strcmp(argv[1], "Binary") == 0 ? Binary : Ascii;
    std::FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 3;
    std::vector<char> data;
    char tmp[4096];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) data.insert(data.end(), tmp, tmp + n);
    std::fclose(f);
    FileSink sink(&st);
    if (sink.open(argv[3]) != 0) return 4;
    size_t off = 0;
    for (int a = 4; a < argc; a++) {
        const size_t len = (size_t)std::strtoull(argv[a], nullptr, 10);
        if (off + len > data.size()) return 5;
        Buffer b(len + 64);
        std::memcpy(b.getData(), data.data() + off, len);
        b.setLength(len);
        const size_t w = sink.write(&b);
        std::printf("%zu\n", w);
        off += len;
    }
    sink.close();
    return 0;
}
