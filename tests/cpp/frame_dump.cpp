// frame_dump.cpp -- writes the frames the CLI reads from a PNG, a PNG sequence
// or a GIF (speckle-stacker_amd/cli/frame_source.h), for one colour mask, to a
// raw file, for comparison with an independent decoder in the tests.
//   smip_frame_dump <in> <mask> <out.raw>   prints "<frames> <cols> <rows> <dtype>"
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "cli/frame_source.h"

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: smip_frame_dump <in> <mask> <out.raw>\n");
        return 2;
    }
    try {
        const std::string in = argv[1];
        const int mask = std::atoi(argv[2]);
        std::unique_ptr<smip_cli::FrameSource> src;
        if (in.size() >= 4 && (in.compare(in.size() - 4, 4, ".png") == 0 || in.compare(in.size() - 4, 4, ".PNG") == 0))
            src = std::make_unique<smip_cli::PngSource>(in, mask);
        else
            src = std::make_unique<smip_cli::GifSource>(in, mask);
        std::vector<uint8_t> buf(src->frame_bytes());
        FILE* f = std::fopen(argv[3], "wb");
        if (!f)
            return 1;
        for (uint64_t i = 0; i < src->nframes(); ++i) {
            src->read(i, 1, buf.data());
            std::fwrite(buf.data(), 1, buf.size(), f);
        }
        std::fclose(f);
        std::printf("%llu %llu %llu %d\n", (unsigned long long)src->nframes(), (unsigned long long)src->ncols(),
            (unsigned long long)src->nrows(), (int)src->dtype());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
