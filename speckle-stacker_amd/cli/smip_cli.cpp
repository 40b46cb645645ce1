// smip_cli -- the speckle-masking command line of the reference
// (smip-cli/main.cpp:40-401) on the MI355X engine.
//
// Same options, defaults and outputs: -n (frames, default 10000), -r, -p (phase
// radius, default 40), -b (bispectrum extent, default 20), -c r|g|b|i, -k, -s,
// -v, -h/-?, --version, --calcsum/--no-calcsum, --specklemasking/
// --no-specklemasking.  As in the reference, -r, -k, -s and the calcsum /
// specklemasking switches are accepted and have no effect (the reference frame
// is always the first one), and -f, listed by the usage text but not parsed,
// prints the usage.  Outputs in the working directory: bispectrum.dat (the
// normalised bispectrum as complex<float>, include/bispectrum.h:411-437) and the
// nine PNG images of smip-cli/main.cpp:380-388.
//
// Input: an animated GIF, a NumPy .npy stack (uint8 / uint16 / float32,
// (frames, rows, cols)) or, with --raw COLSxROWS:u8|u16|f32, a headerless
// stack.  Engine options beyond the reference: --precision mixed|fp64 (default
// mixed: FP32 triple products per batch folded into the FP64 bispectrum),
// --batch N (frames per FP64 fold), --device N, --outdir DIR, --stats (one JSON
// line of stage timings on stdout).
#include <getopt.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "frame_source.h"
#include "png_out.h"
#include "smip_mi355x.h"

using namespace smip_cli;

namespace {

int g_level = 1; // 0 notice, 1 info, 2 debug (the reference starts at info)

struct Log {
    int level;
    std::ostringstream s;
    explicit Log(int l)
        : level(l)
    {
    }
    ~Log()
    {
        if (level <= g_level)
            std::cerr << s.str() << std::endl;
    }
    template <typename T>
    Log& operator<<(const T& v)
    {
        s << v;
        return *this;
    }
};
Log notice() { return Log(0); }
Log info() { return Log(1); }
Log debug() { return Log(2); }

void Usage(const char* progname)
{
    std::cout << "   Usage :  " << progname << " [nrpbcvh?] <source root>\n"
              << "    available options:\n"
              << "     -n   --nrframes    <pics>    :   process at most number of <pics> frames\n"
              << "                                      default : all frames\n"
              << "     -r   --refframe    <index>   :   index of reference frame (default : first found)\n"
              << "     -p   --recoradius  <pixels>  :   radius of phase reconstruction (default : 2 * bispectrum extent)\n"
              << "     -b   --bdepth       <pixels>  :   bispectrum extent (3rd and 4th dimension) (default : 20)\n"
              << "     -c   --channel     <r|g|b|i> :   color channel (default: i)\n"
              << "          --calcsum               :   calculate picture sum and shifted sum (default)\n"
              << "          --no-calcsum            :   do not calculate picture sum and shifted sum\n"
              << "          --specklemasking        :   perform speckle masking (default)\n"
              << "          --no-specklemasking     :   do not perform speckle masking\n"
              << "     -k   --croppos <l:t>         :   fixed crop box position (default: 0:0)\n"
              << "     -s   --cropsize <w:h>        :   crop box size (default: total image)\n"
              << "     -f   --follow                :   follow object defined by the ref frame:  shifts the crop box\n"
              << "                                       for each frame such that the object defined by the reference frame\n"
              << "                                       is always centered. if no crop box position is specified\n"
              << "                                       by --croppos, the brightest object within the ref. image is\n"
              << "                                       automatically selected and the crop box centered around it.\n"
              << "     -v   --verbose               :   increase verbosity level\n"
              << "          --version               :   display version and exit\n"
              << "     -h -?  --help                :   help (this screen)\n"
              << "    engine options:\n"
              << "          --precision mixed|fp64  :   triple-product precision (default: mixed)\n"
              << "          --batch <frames>        :   frames per FP64 fold in mixed precision (default: 512)\n"
              << "          --raw <cols>x<rows>:<u8|u16|f32> : read a headerless frame stack\n"
              << "          --device <index>        :   HIP device (default: 0)\n"
              << "          --outdir <dir>          :   directory of the output files (default: .)\n"
              << "          --stats                 :   print stage timings as one JSON line\n"
              << std::endl;
}

void check(int rc, const char* what)
{
    if (rc != SMIP_OK)
        throw std::runtime_error(std::string(what) + ": " + smip_last_error());
}

int32_t parse_dtype(const std::string& s)
{
    if (s == "u8")
        return SMIP_DTYPE_U8;
    if (s == "u16")
        return SMIP_DTYPE_U16;
    if (s == "f32")
        return SMIP_DTYPE_F32;
    throw std::invalid_argument("unknown sample type " + s);
}

std::unique_ptr<FrameSource> open_source(const std::string& path, const std::string& raw, int colour)
{
    if (!raw.empty()) {
        unsigned long X = 0, Y = 0;
        char t[8] = {};
        if (std::sscanf(raw.c_str(), "%lux%lu:%7s", &X, &Y, t) != 3)
            throw std::invalid_argument("--raw expects <cols>x<rows>:<u8|u16|f32>");
        return std::make_unique<RawSource>(path, 0, X, Y, parse_dtype(t));
    }
    const auto dot = path.rfind('.');
    std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
    std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
    if (ext == "npy")
        return open_npy(path);
    if (ext == "png")
        return std::make_unique<PngSource>(path, colour);
    return std::make_unique<GifSource>(path, colour);
}

} // namespace

int main(int argc, char* argv[])
{
    const char* progname = argv[0];
    int32_t vmaj = 0, vmin = 0, vpat = 0;
    smip_version(&vmaj, &vmin, &vpat);

    std::size_t max_frames { 10'000 };
    std::size_t ref_frame { 0 };
    std::size_t bispectrum_depth { 20 };
    std::size_t reco_radius = bispectrum_depth * 2;
    int colour = 7; // i
    int swSpeckleMasking { 1 };
    int swCalcSum { 1 };
    int swShowVersion { 0 };
    int swStats { 0 };
    std::size_t verbose { 0 };
    std::string precision = "mixed", raw, outdir = ".";
    int batch = 0, device = 0;

    enum { kPrecision = 1000, kBatch, kRaw, kDevice, kOutdir };
    static struct option long_options[] = {
        { "verbose", no_argument, 0, 'v' },
        { "bdepth", required_argument, 0, 'b' },
        { "recoradius", required_argument, 0, 'p' },
        { "refframe", required_argument, 0, 'r' },
        { "nrframes", required_argument, 0, 'n' },
        { "channel", required_argument, 0, 'c' },
        { "croppos", required_argument, 0, 'k' },
        { "cropsize", required_argument, 0, 's' },
        { "help", no_argument, 0, 'h' },
        { "version", no_argument, &swShowVersion, 1 },
        { "no-calcsum", no_argument, &swCalcSum, 0 },
        { "calcsum", no_argument, &swCalcSum, 1 },
        { "no-specklemasking", no_argument, &swSpeckleMasking, 0 },
        { "specklemasking", no_argument, &swSpeckleMasking, 1 },
        { "precision", required_argument, 0, kPrecision },
        { "batch", required_argument, 0, kBatch },
        { "raw", required_argument, 0, kRaw },
        { "device", required_argument, 0, kDevice },
        { "outdir", required_argument, 0, kOutdir },
        { "stats", no_argument, &swStats, 1 },
        { 0, 0, 0, 0 },
    };
    try {
        for (int ch = 0; ch != -1;) {
            int option_index = 0;
            ch = getopt_long(argc, argv, "vn:r:p:b:c:h?k:s:", long_options, &option_index);
            std::istringstream istr;
            switch (ch) {
            case 'v':
                verbose++;
                break;
            case 'n':
                max_frames = std::strtoul(optarg, nullptr, 10);
                break;
            case 'r':
                ref_frame = std::strtoul(optarg, nullptr, 10);
                break;
            case 'p':
                reco_radius = std::strtoul(optarg, nullptr, 10);
                break;
            case 'b':
                bispectrum_depth = std::strtoul(optarg, nullptr, 10);
                break;
            case 'k':
            case 's': {
                istr.str(std::string(optarg));
                int a = -1, b = -1;
                char c;
                istr >> a >> c >> b;
                if (ch == 'k' ? (a < 0 || b < 0) : (a <= 0 || b <= 0))
                    throw std::range_error(ch == 'k' ? "invalid crop box pos arguments"
                                                     : "invalid crop box size arguments");
                break;
            }
            case 'c':
                for (const char c : std::string(optarg)) {
                    if (c == 'r')
                        colour = 4;
                    else if (c == 'g')
                        colour = 2;
                    else if (c == 'b')
                        colour = 1;
                    else if (c == 'i')
                        colour = 7;
                }
                break;
            case kPrecision:
                precision = optarg;
                if (precision != "mixed" && precision != "fp64")
                    throw std::invalid_argument("--precision must be mixed or fp64");
                break;
            case kBatch:
                batch = std::atoi(optarg);
                break;
            case kRaw:
                raw = optarg;
                break;
            case kDevice:
                device = std::atoi(optarg);
                break;
            case kOutdir:
                outdir = optarg;
                break;
            case 'h':
            case '?':
                Usage(progname);
                return 0;
            default:
                break;
            }
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    (void)ref_frame;
    (void)swCalcSum;
    (void)swSpeckleMasking;

    if (swShowVersion) {
        std::cout << "Speckle Masking Image Processing v" << vmaj << "." << vmin << "." << vpat << std::endl;
        return 0;
    }
    g_level = verbose == 0 ? 1 : (verbose == 1 ? 1 : 2);
    argc -= optind;
    argv += optind;
    if (argc != 1) {
        Usage(progname);
        return 0;
    }
    const std::string filename(*argv);
    info() << "Speckle Masking Image Processing CLI v" << vmaj << "." << vmin << "." << vpat << " (MI355X engine)";

    smip_ctx* ctx = nullptr;
    smip_job* job = nullptr;
    try {
        auto src = open_source(filename, raw, colour);
        info() << "opening video file " << filename;
        const uint64_t nframes = std::min<uint64_t>(src->nframes(), max_frames);
        info() << "opened video file: " << src->nframes() << " frames";
        notice() << "using " << nframes << "/" << src->nframes() << " frames";
        if (nframes == 0)
            throw std::runtime_error("no frames to process");
        const uint64_t X = src->ncols(), Y = src->nrows();
        debug() << "creating bispectrum with size [" << X << " " << Y << " " << bispectrum_depth << " "
                << bispectrum_depth << "]";

        check(smip_ctx_create(device, &ctx), "context");
        smip_job_params jp {};
        jp.ncols = X;
        jp.nrows = Y;
        jp.bdepth = bispectrum_depth;
        jp.reco_radius = static_cast<double>(reco_radius);
        jp.frame_dtype = src->dtype();
        jp.precision = precision == "fp64" ? SMIP_PRECISION_FP64 : SMIP_PRECISION_MIXED;
        jp.saa = 1;
        jp.frames_per_batch = batch;
        check(smip_job_create(ctx, &jp, &job), "job");

        info() << "creating sum, power spectra and accumulating bispectrum of all frames";
        const uint64_t chunk = 512;
        std::vector<uint8_t> buf(chunk * src->frame_bytes());
        for (uint64_t f = 0; f < nframes; f += chunk) {
            const uint64_t n = std::min(chunk, nframes - f);
            debug() << "reading frames " << f + 1 << ".." << f + n << "/" << nframes;
            src->read(f, n, buf.data());
            check(smip_job_add_frames_host(job, buf.data(), n), "accumulate");
        }
        info() << "normalizing, reconstructing fourier phases from bispectrum, synthesizing the image";
        check(smip_job_finalize(job), "finalize");

        const uint64_t np = X * Y;
        std::vector<double> sum(np), power(np), phases(2 * np), cons(np), image(2 * np);
        std::vector<uint8_t> flags(np);
        check(smip_job_get_sum_image(job, sum.data()), "sum image");
        check(smip_job_get_powerspec(job, power.data()), "power spectrum");
        check(smip_job_get_phases(job, phases.data(), flags.data(), cons.data()), "phases");
        check(smip_job_get_image(job, image.data()), "image");

        notice() << "writing bispectrum to file 'bispectrum.dat'";
        {
            smip_bispec* b64 = nullptr; // the job's own accumulator, borrowed
            smip_bispec* b32 = nullptr;
            check(smip_job_bispectrum(job, &b64), "bispectrum");
            const uint64_t dims[4] = { X, Y, bispectrum_depth, bispectrum_depth };
            check(smip_bispec_create(ctx, dims, SMIP_DTYPE_C64, &b32), "bispectrum");
            int rc = smip_bispec_assign(b32, b64);
            if (rc == SMIP_OK)
                rc = smip_bispec_write_file(b32, (outdir + "/bispectrum.dat").c_str());
            if (rc != SMIP_OK) {
                const std::string err = smip_last_error();
                smip_bispec_destroy(b32);
                throw std::runtime_error("bispectrum.dat: " + err);
            }
            smip_bispec_destroy(b32);
        }

        // phases *= Hann(X, Y, 2p) for the phase images (smip-cli/main.cpp:328-330)
        std::vector<double> win(np);
        check(smip_general_hamming(X, Y, 2.0 * static_cast<double>(reco_radius), 0.5, win.data()), "window");
        std::vector<std::complex<double>> wph(np);
        for (uint64_t k = 0; k < np; ++k)
            wph[k] = std::complex<double>(phases[2 * k], phases[2 * k + 1]) * std::complex<double>(win[k], 0.0);
        // power spectrum scaled to a maximum of 1000 (smip-cli/main.cpp:361-366)
        const double pmax = *std::max_element(power.begin(), power.end(),
            [](double a, double b) { return std::fabs(a) < std::fabs(b); });
        const double pnorm = std::fabs(pmax) / 1000.;
        const std::string d = outdir + "/";
        const auto sum_at = [&](int64_t k) { return std::fabs(sum[k] / 255.); };
        const auto phase_at = [&](int64_t k) { return (std::arg(wph[k]) + M_PI) / (2.0 * M_PI); };
        const auto cons_at = [&](int64_t k) { return cons[k]; };
        const auto power_at = [&](int64_t k) { return std::fabs(power[k] / pnorm); };
        const auto image_at = [&](int64_t k) { return std::abs(std::complex<double>(image[2 * k], image[2 * k + 1])); };
        save_image(d + "sum_image_falsecolor.png", X, Y, true, false, sum_at);
        save_image(d + "sum_image.png", X, Y, false, false, sum_at);
        save_image(d + "phases_falsecolor.png", X, Y, true, true, phase_at);
        save_image(d + "phases.png", X, Y, false, true, phase_at);
        save_image(d + "phasecons.png", X, Y, true, true, cons_at);
        save_image(d + "powerspec_falsecolor.png", X, Y, true, true, power_at);
        save_image(d + "powerspec.png", X, Y, false, true, power_at);
        save_image(d + "reco_image_falsecolor.png", X, Y, true, true, image_at);
        save_image(d + "reco_image.png", X, Y, false, true, image_at);

        smip_job_stats st {};
        check(smip_job_get_stats(job, &st), "stats");
        uint64_t nflag = 0;
        for (auto f : flags)
            nflag += f != 0;
        info() << "frames " << st.frames << ", accumulate " << st.accumulate_ms << " ms, shift-and-add " << st.saa_ms
               << " ms, phases " << st.reco_ms << " ms (" << nflag << " recovered), synthesis " << st.synth_ms << " ms";
        if (swStats)
            std::printf("{\"frames\": %llu, \"ncols\": %llu, \"nrows\": %llu, \"bdepth\": %zu, \"reco_radius\": %zu, "
                        "\"precision\": \"%s\", \"accumulate_ms\": %.3f, \"saa_ms\": %.3f, \"reco_ms\": %.3f, "
                        "\"synth_ms\": %.3f, \"recovered\": %llu}\n",
                (unsigned long long)st.frames, (unsigned long long)X, (unsigned long long)Y, bispectrum_depth,
                reco_radius, precision.c_str(), st.accumulate_ms, st.saa_ms, st.reco_ms, st.synth_ms,
                (unsigned long long)nflag);
        smip_job_destroy(job);
        smip_ctx_destroy(ctx);
    } catch (const std::exception& e) {
        std::cerr << "smip_cli: " << e.what() << std::endl;
        if (job)
            smip_job_destroy(job);
        if (ctx)
            smip_ctx_destroy(ctx);
        return 1;
    }
    return 0;
}
