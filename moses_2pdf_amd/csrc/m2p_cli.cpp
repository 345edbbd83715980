// m2p_cli.cpp -- MoSES_2PDF_landslide / MoSES_2PDF_inflow executables.
//
// Same contract as the reference's main (MoSES_2PDF_landslide/main.cu:46-58): read
// ./par_List, write ./result2Pb/, print the banner, the parameters and
// "Time used on GPU: %.2f sec".  Optional argv[1] is an overlay string
// ("cfl=0.4;rho_ratio=0.37"), argv[2] a result directory.
#include <cstdio>

#include "moses2pdf.h"

#ifndef M2P_CLI_MODE
#define M2P_CLI_MODE 0
#endif

int main(int argc, char** argv) {
  const int mode = M2P_CLI_MODE;
  const char* overlay = argc > 1 ? argv[1] : nullptr;
  const char* outdir = argc > 2 ? argv[2] : "./result2Pb/";
  std::fprintf(stdout, "\n\t***********************************************************************************\n");
  std::fprintf(stdout, mode == 0 ? "\t\t* * * * *     MoSES_2PDF:  landslide (debris flow) mode     * * * * * \n"
                                 : "\t\t* * * * *     MoSES_2PDF:  inflow (debris flow) mode     * * * * * \n");
  std::fprintf(stdout, "\t-----------------------------------------------------------------------------------\n");
  std::fprintf(stdout, "\t\t2D Central Scheme  (adNOC) Mixture Code, MI355X HIP engine: %s mode\n",
               mode == 0 ? "landslide" : "inflow");
  std::fprintf(stdout, "\t***********************************************************************************\n\n");
  m2p_ctx* ctx = nullptr;
  int rc = m2p_open("./par_List", mode, overlay, &ctx);
  if (rc) {
    std::printf("\n\t----------------------------------\n\t   Error: %s\n\t----------------------------------\n",
                m2p_last_error(ctx));
    m2p_close(ctx);
    return 1;
  }
  int nx, ny;
  m2p_dims(ctx, &nx, &ny, nullptr, nullptr);
  std::fprintf(stdout, "\t\tData points          : %d,%d\n", nx, ny);
  double wall = 0;
  rc = m2p_run(ctx, outdir, &wall);
  if (rc) {
    std::printf("\n\t----------------------------------\n\t   Error: %s\n\t----------------------------------\n",
                m2p_last_error(ctx));
    m2p_close(ctx);
    return 1;
  }
  m2p_close(ctx);
  std::printf("Time used on GPU: %.2f sec\n", wall);
  return 0;
}
