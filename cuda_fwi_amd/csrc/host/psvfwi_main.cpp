// psvfwi_main.cpp -- command-line driver, as the reference's
// CUDA_week2/main.cu:6-25: `psvfwi_main config.ini` builds the model from the
// INI file and runs fwi_simulate when [FDM] full_waveform_inversion is set,
// forward_simulate otherwise.  Errors are reported with exit status 1.
#include <exception>
#include <iostream>

#include "fd_host.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " configuration.ini" << std::endl;
        return 2;
    }
    try {
        FDWaveModel model(argv[1]);
        if (model.fwinv) model.fwi_simulate();
        else model.forward_simulate();
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "PROGRAM COMPLETE." << std::endl;
    return 0;
}
