// phyloacc-st-gpu: the drop-in command line (pa_cli_main in the engine library).
extern "C" int pa_cli_main(int argc, char **argv);

int main(int argc, char **argv) { return pa_cli_main(argc, argv); }
