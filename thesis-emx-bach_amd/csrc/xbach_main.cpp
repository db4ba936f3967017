// xbach_main.cpp — the command-line driver: the reference's main
// (src/main.cpp:16-159) and argument parsing (src/argsUtil.cpp:23-81) over
// libxbach.
//
//   xbach -t <template> -s <science> [-o diff.fits] [-op out/] [-ip res/]
//         [-sx 10] [-sy 10] [-sss mp|cl] [-p 0] [-d 0] [-v] [-vt]
//
// Input files are <ip>/<name>; outputs are <op><o> (the convolved image) and
// <op>sub.fits (the difference image), paths concatenated exactly as the
// reference does (src/main.cpp:35, include/datatypeUtil.h:160-169).  With
// -vt the nine "<Stage> took N ms" lines are printed in the reference's order
// (Ini SSS CMV CD KSC Conv Sub Fin BACH), each stage timed to its completion.
// The difference image carries HOTPANTS's quality cards (KSUM00, X2NRM00 ...;
// tests/test01_sub.fits) computed by xb_get_diagnostics.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xbach.h"

namespace {

struct Arguments {  // include/argsUtil.h:7-57 (the fields this path reads)
  std::string templateName, scienceName;
  std::string outName = "diff.fits";
  std::string inputPath = "res/";
  std::string outPath = "out/";
  std::string sssMode = "mp";
  int platform = 0, device = 0;
  int stampsx = 10, stampsy = 10;
  bool verbose = false, verboseTime = false;
};

const char* option(int argc, const char* argv[], const std::string& o) {
  const char** end = argv + argc;
  const char** it = std::find(argv, end, o);
  if (it != end && ++it != end) return *it;
  return nullptr;
}

bool has(int argc, const char* argv[], const std::string& o) {
  return std::find(argv, argv + argc, o) != argv + argc;
}

// src/argsUtil.cpp:23-81: same options, same required-argument messages.
void getArguments(int argc, const char* argv[], Arguments& a) {
  auto str = [&](const char* o, std::string& dst) {
    if (has(argc, argv, o)) {
      const char* v = option(argc, argv, o);
      if (!v) throw std::invalid_argument(std::string("missing value for ") + o);
      dst = v;
    }
  };
  auto num = [&](const char* o, int& dst) {
    if (has(argc, argv, o)) {
      const char* v = option(argc, argv, o);
      if (!v) throw std::invalid_argument(std::string("missing value for ") + o);
      std::stringstream s{v};
      s >> dst;
    }
  };
  str("-o", a.outName);
  str("-op", a.outPath);
  str("-ip", a.inputPath);
  num("-sx", a.stampsx);
  num("-sy", a.stampsy);
  if (has(argc, argv, "-sss")) {
    str("-sss", a.sssMode);
    if (a.sssMode != "mp" && a.sssMode != "cl")
      throw std::invalid_argument("sss mode should be mp or cl");
  }
  a.verbose = has(argc, argv, "-v");
  a.verboseTime = has(argc, argv, "-vt");
  if (has(argc, argv, "-p")) {
    const char* v = option(argc, argv, "-p");
    a.platform = v ? std::atoi(v) : 0;
  }
  if (has(argc, argv, "-d")) {
    const char* v = option(argc, argv, "-d");
    a.device = v ? std::atoi(v) : 0;
  }
  if (has(argc, argv, "-t")) str("-t", a.templateName);
  else throw std::invalid_argument("Template file Input is required!");
  if (has(argc, argv, "-s")) str("-s", a.scienceName);
  else throw std::invalid_argument("Science file input is required!");
}

using Clock = std::chrono::steady_clock;
long long ms(Clock::time_point b, Clock::time_point a) {
  return std::chrono::duration_cast<std::chrono::milliseconds>(b - a).count();
}

// A string-valued header card laid out as HOTPANTS writes them
// (tests/test01_sub.fits): KEY     = 'value   '           / comment
std::string card(const char* key, const std::string& value, const char* comment) {
  std::string v = value;
  if (v.size() < 8) v.resize(8, ' ');
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%-8s= %-20s / %s", key, ("'" + v + "'").c_str(), comment);
  std::string c = buf;
  c.resize(80, ' ');
  return c;
}

std::string num(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.4f", v);
  return b;
}

// HOTPANTS's quality cards for the difference image (xb_get_diagnostics).
std::string diagCards(const xb_diag& d, int w, int h) {
  std::string c;
  c += card("SOFTNAME", "X-BACH MI355X", "The software that differenced this image");
  c += card("REGION00", "[1:" + std::to_string(w) + ",1:" + std::to_string(h) + "]",
            "Region of the image");
  c += card("CONVOL00", d.convTemplate ? "TEMPLATE" : "IMAGE", "Image that was convolved");
  c += card("KSUM00", num(d.kernSum), "Kernel Sum");
  c += card("SSSIG00", num(d.sssig), "Average Figure of Merit across Stamps");
  c += card("SSSCAT00", num(d.ssscat), "Stdev in Figure of Merit");
  c += card("FSIG00", num(d.fsig), "Final SSSIG");
  c += card("FSCAT00", num(d.fscat), "Final SSSCAT");
  c += card("X2NRM00", num(d.x2nrm), "1/N * SUM (diff/noise)^2");
  c += card("NX2NRM00", std::to_string(d.nx2nrm), "Number of pixels in X2NRM");
  c += card("DMEAN00", num(d.dmean), "Mean of diff image; good pixels");
  c += card("DSIGE00", num(d.dsige), "Stdev of diff image; good pixels");
  c += card("DSIG00", num(d.dsig), "Mean of noise image; good pixels");
  c += card("DMEANO00", num(d.dmeano), "Mean of diff image; OK pixels");
  c += card("DSIGEO00", num(d.dsigeo), "Stdev of diff image; OK pixels");
  c += card("DSIGO00", num(d.dsigo), "Mean of noise image; OK pixels");
  c += card("PHOTNORM", d.normalizeTemplate ? "t" : "i", "Direction of photometric normalization");
  char b[96];
  std::snprintf(b, sizeof(b), "%-8s= %20d / %s", "NREGION", 1, "Number of independent regions");
  std::string r = b;
  r.resize(80, ' ');
  c += r;
  std::snprintf(b, sizeof(b), "%-8s= %20s / %s", "MASKVAL", "1.0E-30", "Value of Masked Pixels");
  r = b;
  r.resize(80, ' ');
  c += r;
  return c;
}

int fail(xb_ctx* ctx, int rc, const char* what) {
  std::cout << what << ": " << (ctx ? xb_last_error(ctx) : "") << " (status " << rc << ")"
            << std::endl;
  if (ctx) xb_ctx_destroy(ctx);
  return 1;
}

}  // namespace

int main(int argc, const char* argv[]) {
  const auto p1 = Clock::now();
  Arguments args;
  try {
    std::cout << "Reading in arguments..." << std::endl;
    getArguments(argc, argv, args);
  } catch (const std::invalid_argument& err) {
    std::cout << err.what() << '\n';
    return 1;
  }
  std::cout << "\nReading in images..." << std::endl;
  const std::string inDir = args.inputPath + "/";
  if (args.verbose)
    std::cout << "template image name: " << args.templateName
              << ", science image name: " << args.scienceName << std::endl;

  std::cout << "\nSetting up HIP..." << std::endl;
  xb_ctx* ctx = nullptr;
  int rc = xb_ctx_create(args.device, 0, &ctx);
  if (rc != XB_OK) {
    std::cout << "No usable gfx950 GPU at device " << args.device << " (status " << rc << ")"
              << std::endl;
    return 1;
  }
  if (args.verbose) {
    char name[256];
    xb_device_name(ctx, name, sizeof(name));
    std::cout << "device: " << name << std::endl;
  }

  // init (src/bach.cpp:14-42): read both images, size check, masks
  char msg[512] = {0};
  int w = 0, h = 0, ws = 0, hs = 0;
  const std::string tPath = inDir + args.templateName, sPath = inDir + args.scienceName;
  if (xb_fits_read(tPath.c_str(), &w, &h, nullptr, 0, msg, sizeof(msg)) != XB_OK ||
      xb_fits_read(sPath.c_str(), &ws, &hs, nullptr, 0, msg, sizeof(msg)) != XB_OK) {
    std::cout << msg << std::endl;
    xb_ctx_destroy(ctx);
    return 1;
  }
  if (w != ws || h != hs) {
    std::cout << "Template image and science image must be the same size!" << std::endl;
    xb_ctx_destroy(ctx);
    return 1;
  }
  const size_t n = (size_t)w * h;
  std::vector<float> tmpl(n), sci(n);
  if (xb_fits_read(tPath.c_str(), &w, &h, tmpl.data(), n, msg, sizeof(msg)) != XB_OK ||
      xb_fits_read(sPath.c_str(), &w, &h, sci.data(), n, msg, sizeof(msg)) != XB_OK) {
    std::cout << msg << std::endl;
    xb_ctx_destroy(ctx);
    return 1;
  }
  xb_params p;
  xb_params_default(&p);
  p.stampsx = args.stampsx;
  p.stampsy = args.stampsy;
  p.verbose = args.verbose ? 1 : 0;
  if ((rc = xb_init(ctx, &p, w, h, tmpl.data(), sci.data())) != XB_OK || (rc = xb_sync(ctx)))
    return fail(ctx, rc, "init");
  const auto p2 = Clock::now();
  if (args.verboseTime) std::cout << "Ini took " << ms(p2, p1) << " ms" << std::endl;

  const auto p3 = Clock::now();
  if ((rc = xb_sss(ctx)) != XB_OK || (rc = xb_sync(ctx))) {
    if (rc == XB_ERR_NO_SUBSTAMPS) {  // src/bach.cpp:166-169
      std::cout << "No substamps found" << std::endl;
      xb_ctx_destroy(ctx);
      return 1;
    }
    return fail(ctx, rc, "sss");
  }
  const auto p4 = Clock::now();
  if (args.verboseTime) std::cout << "SSS took " << ms(p4, p3) << " ms" << std::endl;
  std::cout << std::endl;

  const auto p5 = Clock::now();
  if ((rc = xb_cmv(ctx)) != XB_OK || (rc = xb_sync(ctx))) return fail(ctx, rc, "cmv");
  const auto p6 = Clock::now();
  if (args.verboseTime) std::cout << "CMV took " << ms(p6, p5) << " ms" << std::endl;

  const auto p7 = Clock::now();
  int convTemplate = 1;
  if ((rc = xb_cd(ctx, &convTemplate)) != XB_OK || (rc = xb_sync(ctx))) return fail(ctx, rc, "cd");
  const auto p8 = Clock::now();
  if (args.verboseTime) std::cout << "CD took " << ms(p8, p7) << " ms" << std::endl;
  if (args.verbose)
    std::cout << (convTemplate ? "convolving the template" : "convolving the science image")
              << std::endl;

  const auto p9 = Clock::now();
  if ((rc = xb_ksc(ctx)) != XB_OK || (rc = xb_sync(ctx))) return fail(ctx, rc, "ksc");
  const auto p10 = Clock::now();
  if (args.verboseTime) std::cout << "KSC took " << ms(p10, p9) << " ms" << std::endl;

  const auto p11 = Clock::now();
  double kernSum = 0.0;
  if ((rc = xb_conv(ctx, &kernSum)) != XB_OK || (rc = xb_sync(ctx))) return fail(ctx, rc, "conv");
  const auto p12 = Clock::now();
  if (args.verboseTime) std::cout << "Conv took " << ms(p12, p11) << " ms" << std::endl;
  if (args.verbose) std::cout << "kernel sum " << kernSum << std::endl;

  const auto p13 = Clock::now();
  if ((rc = xb_sub(ctx)) != XB_OK || (rc = xb_sync(ctx))) return fail(ctx, rc, "sub");
  const auto p14 = Clock::now();
  if (args.verboseTime) std::cout << "Sub took " << ms(p14, p13) << " ms" << std::endl;

  // fin (src/bach.cpp:442-447)
  const auto p15 = Clock::now();
  std::cout << "\nWriting output..." << std::endl;
  std::vector<float> conv(n), diff(n);
  if ((rc = xb_get_conv(ctx, conv.data())) != XB_OK || (rc = xb_get_diff(ctx, diff.data())))
    return fail(ctx, rc, "fin");
  const std::string convOut = args.outPath + args.outName, diffOut = args.outPath + "sub.fits";
  xb_diag diag;
  if ((rc = xb_get_diagnostics(ctx, &diag)) != XB_OK) return fail(ctx, rc, "fin");
  const std::string cards = diagCards(diag, w, h);
  if (xb_fits_write(convOut.c_str(), w, h, conv.data(), msg, sizeof(msg)) != XB_OK ||
      xb_fits_write_cards(diffOut.c_str(), w, h, diff.data(), cards.data(),
                          (int)(cards.size() / 80), msg, sizeof(msg)) != XB_OK) {
    std::cout << msg << std::endl;
    xb_ctx_destroy(ctx);
    return 1;
  }
  const auto p16 = Clock::now();
  if (args.verboseTime) std::cout << "Fin took " << ms(p16, p15) << " ms" << std::endl;
  xb_ctx_destroy(ctx);
  std::cout << "\nBACH finished." << std::endl;
  if (args.verboseTime) std::cout << "BACH took " << ms(p16, p1) << " ms" << std::endl;
  return 0;
}
