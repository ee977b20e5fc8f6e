// ct_forward_projection -- command-line driver (the reference's
// Main_Forward_Projection_V1.1.cpp:18-41): reads Config/ScanSetting.ini and
// Config/OutPutFolder.ini under the directory given as the first argument
// (default: the working directory) and runs the configured scan.
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "ForwardProjection.h"

int main(int argc, char **argv)
{
    ForwardProjection fp;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--device") && i + 1 < argc)
            fp.device = std::atoi(argv[++i]);
        else
            fp.rootDir = argv[i];
    }
    try {
        fp.run();
    } catch (const ctfp::Error &e) {
        std::cerr << "ct_forward_projection: " << e.what() << "\n";
        return e.code < 0 ? -e.code : 1;
    } catch (const std::exception &e) {
        std::cerr << "ct_forward_projection: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
