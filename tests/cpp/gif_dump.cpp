// gif_dump.cpp -- writes the composited R,G,B frames the CLI decodes from a GIF
// (speckle-stacker_amd/cli/frame_source.h) to a raw file, for comparison with
// an independent decoder in the tests.
//   smip_gif_dump <in.gif> <out.rgb>   prints "<frames> <cols> <rows>"
#include <cstdio>

#include "cli/frame_source.h"

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: smip_gif_dump <in.gif> <out.rgb>\n");
        return 2;
    }
    try {
        smip_cli::GifSource g(argv[1], 7);
        FILE* f = std::fopen(argv[2], "wb");
        if (!f)
            return 1;
        for (size_t i = 0; i < g.nframes(); ++i)
            std::fwrite(g.rgb_frame(i).data(), 1, g.rgb_frame(i).size(), f);
        std::fclose(f);
        std::printf("%llu %llu %llu\n", (unsigned long long)g.nframes(), (unsigned long long)g.ncols(),
            (unsigned long long)g.nrows());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
