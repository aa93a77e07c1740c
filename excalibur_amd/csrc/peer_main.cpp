// Peer executable: the reference's main.cpp command line (main.cpp:16-33, 58-104) in front of the
// GPU DecentralizedClient, plus long options of its own (--form, --refine-iters, --dump,
// --recv-timeout, --repeat, --profile, --ralha-loop, --device-exchange).
#include <getopt.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

#include "client.hpp"

int main(int argc, char** argv) {
  int m = 0, n = 0, port = 0, topk = -1, mode = 1, num_client = 1, cid = 0, opt = 3;
  std::string listen = "127.0.0.1", dpath = ".", logdir = ".";
  bool evaluate = false, is_memmap = false;
  unsigned int seed = 0;
  ClientOptions options;
  static struct option long_options[] = {
      {"listen", required_argument, nullptr, 'l'}, {"port", required_argument, nullptr, 'p'},
      {"drow", required_argument, nullptr, 'm'},   {"dcol", required_argument, nullptr, 'n'},
      {"dpath", required_argument, nullptr, 'd'},  {"logdir", required_argument, nullptr, 'g'},
      {"topk", required_argument, nullptr, 'k'},   {"mode", required_argument, nullptr, 'o'},
      {"nclient", required_argument, nullptr, 't'}, {"cid", required_argument, nullptr, 'c'},
      {"evaluate", required_argument, nullptr, 'e'}, {"ismemmap", required_argument, nullptr, 'f'},
      {"seed", required_argument, nullptr, 's'},   {"opt", required_argument, nullptr, 'a'},
      {"help", no_argument, nullptr, 'h'},         {"form", required_argument, nullptr, 1000},
      {"dump", required_argument, nullptr, 1001},  {"recv-timeout", required_argument, nullptr, 1002},
      {"refine-iters", required_argument, nullptr, 1003}, {"repeat", required_argument, nullptr, 1004},
      {"profile", required_argument, nullptr, 1005},      {"ralha-loop", required_argument, nullptr, 1006},
      {"device-exchange", required_argument, nullptr, 1007},
      {nullptr, 0, nullptr, 0}};
  int c, index = 0;
  while ((c = getopt_long(argc, argv, "l:p:m:n:k:o:d:g:t:c:e:f:s:a:h", long_options, &index)) != -1) {
    switch (c) {
      case 'l':
        listen = optarg;
        // the reference's -l falls through into -p (main.cpp:61-64): -p must follow -l
        port = atoi(optarg);
        break;
      case 'p': port = atoi(optarg); break;
      case 'm': m = atoi(optarg); break;
      case 'n': n = atoi(optarg); break;
      case 'k': topk = atoi(optarg); break;
      case 'o': mode = atoi(optarg); break;
      case 'd': dpath = optarg; break;
      case 'g': logdir = optarg; break;
      case 't': num_client = atoi(optarg); break;
      case 'c': cid = atoi(optarg); break;
      case 'e': evaluate = atoi(optarg) == 1; break;
      case 'f': is_memmap = atoi(optarg) == 1; break;
      case 's': seed = (unsigned)atoi(optarg); break;
      case 'a': opt = atoi(optarg); break;
      case 1000: options.form = optarg; break;
      case 1001: options.dump_results = atoi(optarg) == 1; break;
      case 1002: options.recv_timeout_s = atof(optarg); break;
      case 1003: options.refine_iters = atoi(optarg); break;
      case 1004: options.repeat = atoi(optarg); break;
      case 1005: options.profile = atoi(optarg) == 1; break;
      case 1006: options.ralha_one_launch = atoi(optarg) == 1; break;
      case 1007: options.device_exchange = atoi(optarg) == 1; break;
      case 'h':
        printf(" -l --listen base server IP\n -p --port base server port\n -m --drow rows\n -n --dcol local columns\n"
               " -d --dpath data dir\n -g --logdir log dir\n -k --topk top-k (-1 all)\n -o --mode 0..3\n"
               " -t --nclient peers\n -c --cid my id\n -e --evaluate 1 dumps results\n -s --seed\n -a --opt 1|3\n"
               " -f --ismemmap 1 maps the input file\n"
               " --form fp64|mixed  --refine-iters N  --dump 0|1  --recv-timeout seconds\n"
               " --repeat N  --profile 0|1  --ralha-loop 0|1  --device-exchange 0|1\n");
        return 0;
      default: return 2;
    }
  }
  try {
    DecentralizedClient client(cid, num_client, mode, (size_t)topk, (size_t)m, (size_t)n, port, listen, dpath,
                               is_memmap, evaluate, logdir, seed, opt, options);
  } catch (const std::exception& e) {
    fprintf(stderr, "peer %d failed: %s\n", cid, e.what());
    fflush(stderr);
    _exit(3);
  }
  return 0;
}
