// fs_main.cpp -- the `factsieve` command line, drop-in for the reference
// executable (main.cpp:27-399): same options -n -N -p -P -v -s -d -h (plus
// --test and --device), same number syntax (putil.c), same stdout/stderr lines,
// factors.txt and stateA/B.ckp in the working directory, exit(EXIT_FAILURE) on
// any error.  BOINC is not linked: the program runs standalone, and takes its
// GPU from -d, else from an init_data.xml in the working directory
// (<gpu_device_num>, README.md:42-56), else device 0.
#include <getopt.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/factsieve.h"

#define FS_VERSION "2.0-mi355x"

static void help() {
  printf("Welcome to FactSieve, a GPU program to find factors of factorial prime candidates of the form n!+-1\n");
  printf("Note that there are no factors for numbers of the form n!+-1 when p <= n because all primes <= n are factors of n!\n");
  printf("Program usage:\n");
  printf("-n #\t\t\tStart factorial n!+-1\n");
  printf("-N #\t\t\tEnd factorial N!+-1, range [-n, -N) exclusive, 127 <= -n <= n < -N < 2^31\n");
  printf("-p #\t\t\tStarting prime factor p\n");
  printf("-P #\t\t\tEnd prime factor P, range [-p, -P) exclusive, 127 <= -n <= -p <= p < -P < 2^64\n");
  printf("-v #\t\t\tOptional, specify the number of CPU threads used to verify factors.  Default is 2, max is 128.\n");
  printf("-s \t\t\tPerform self test to verify proper operation of the program with the current GPU.\n");
  printf("-d #\t\t\tOptional, GPU device index (default 0).\n");
  printf("--verify-on cpu|gpu|auto\tOptional, where found factors are re-verified (default auto: on the GPU when the\n"
         "\t\t\tfactors need more than 2^24 multiplications, else on the -v CPU threads).\n");
  printf("-h\t\t\tPrint this help\n");
  exit(EXIT_FAILURE);
}

struct Args {
  uint64_t pmin = 0, pmax = 0;
  uint64_t nmin = 0, nmax = 0;
  uint64_t threads = 2;
  uint64_t device = 0;
  bool device_set = false;
  int verify_device = FS_VERIFY_AUTO;
  bool test = false;
};

// parse_option (main.cpp:45-90)
static int parse_option(int opt, const char* arg, Args& a) {
  switch (opt) {
    case 'p': return fs_parse_uint64(&a.pmin, arg, 127, 0xFFFFFFFFFFFFFFFFull - 1);
    case 'P': return fs_parse_uint64(&a.pmax, arg, 128, 0xFFFFFFFFFFFFFFFFull);
    case 'n': return fs_parse_uint64(&a.nmin, arg, 127, 0x7FFFFFFFull - 1);
    case 'N': return fs_parse_uint64(&a.nmax, arg, 128, 0x7FFFFFFFull);
    case 'v': return fs_parse_uint64(&a.threads, arg, 1, 128);
    case 'd':
      a.device_set = arg != nullptr;
      return arg ? fs_parse_uint64(&a.device, arg, 0, 63) : 0;
    case 'V':
      if (!strcmp(arg, "cpu")) a.verify_device = FS_VERIFY_CPU;
      else if (!strcmp(arg, "gpu")) a.verify_device = FS_VERIFY_GPU;
      else if (!strcmp(arg, "auto")) a.verify_device = FS_VERIFY_AUTO;
      else return -1;
      return 0;
    case 's':
      a.test = true;
      fprintf(stderr, "Performing self test.\n");
      printf("Performing self test.\n");
      return 0;
    case 'h':
    case '?':
      help();
  }
  return 0;
}

static void die(const char* what) {
  printf("%s\n", what);
  fprintf(stderr, "%s\n", what);
  exit(EXIT_FAILURE);
}

int main(int argc, char** argv) {
  static const struct option long_opts[] = {
      {"device", optional_argument, 0, 'd'}, {"test", no_argument, 0, 's'},
      {"verify-on", required_argument, 0, 'V'}, {0, 0, 0, 0}};
  Args a;
  fprintf(stderr, "\nFactSieve v%s\n", FS_VERSION);
  printf("\nFactSieve v%s\n", FS_VERSION);
  fprintf(stderr, "Command line: ");
  for (int i = 0; i < argc; i++) fprintf(stderr, "%s ", argv[i]);
  fprintf(stderr, "\n");
  int opt, ind = -1;
  while ((opt = getopt_long(argc, argv, "p:P:n:N:v:sd:h", long_opts, &ind)) != -1) {
    const int st = parse_option(opt, optarg, a);
    if (st == -1) {
      if (ind == -1) {
        printf("%s: invalid argument -%c %s\n", argv[0], opt, optarg);
        fprintf(stderr, "%s: invalid argument -%c %s\n", argv[0], opt, optarg);
      } else {
        printf("%s: invalid argument --%s %s\n", argv[0], long_opts[ind].name, optarg);
        fprintf(stderr, "%s: invalid argument --%s %s\n", argv[0], long_opts[ind].name, optarg);
      }
      exit(EXIT_FAILURE);
    } else if (st == -2) {
      if (ind == -1) {
        printf("%s: out of range argument -%c %s\n", argv[0], opt, optarg);
        fprintf(stderr, "%s: out of range argument -%c %s\n", argv[0], opt, optarg);
      } else {
        printf("%s: out of range argument --%s %s\n", argv[0], long_opts[ind].name, optarg);
        fprintf(stderr, "%s: out of range argument --%s %s\n", argv[0], long_opts[ind].name, optarg);
      }
      exit(EXIT_FAILURE);
    }
    ind = -1;
  }

  if (!a.device_set) {  // boinc_get_opencl_ids, stand-alone form (main.cpp:230-247)
    int dev = 0;
    char type[64];
    const int found = fs_read_init_data("", &dev, type, sizeof type);
    if (found == 1) {
      a.device = (uint64_t)dev;
      printf("init_data.xml: using GPU device %d (%s).\n", dev, type[0] ? type : "gpu_type not given");
      fprintf(stderr, "init_data.xml: using GPU device %d (%s).\n", dev, type[0] ? type : "gpu_type not given");
    } else if (found == 0) {
      printf("init_data.xml not found, using device 0.\n");
    } else {
      printf("%s, using device 0.\n", fs_last_error());
      fprintf(stderr, "%s, using device 0.\n", fs_last_error());
    }
  }
  int rc = fs_init((int)a.device);
  if (rc) die(fs_last_error());
  char name[256];
  uint32_t cus = 0;
  uint64_t total = 0, freeb = 0;
  if (fs_device_info(name, sizeof name, &cus, &total, &freeb) == FS_OK) {
    fprintf(stderr, "GPU Info:\n  Name: \t\t%s\n  Compute Units: \t%u\n  Memory: \t\t%llu MiB\n", name, cus,
            (unsigned long long)(total >> 20));
    printf("GPU Info:\n  Name: \t\t%s\n  Compute Units: \t%u\n  Memory: \t\t%llu MiB\n", name, cus,
           (unsigned long long)(total >> 20));
  }

  fs_options o;
  fs_default_options(&o);
  o.verify_threads = (int)a.threads;
  o.verify_device = a.verify_device;
  o.write_files = 1;
  o.resume = 1;
  o.quiet = 0;
  if (a.test) {
    const int good = fs_self_test(&o, nullptr);
    return good == 5 ? EXIT_SUCCESS : EXIT_FAILURE;
  }
  if (a.pmin == 0 || a.pmax == 0) die("-p and -P arguments are required\nuse -h for help");
  if (a.nmin == 0 || a.nmax == 0) die("-n and -N arguments are required\nuse -h for help");
  fs_result r;
  rc = fs_sieve(a.pmin, a.pmax, (uint32_t)a.nmin, (uint32_t)a.nmax, &o, &r, nullptr, 0, nullptr);
  if (rc) die(fs_last_error());
  return EXIT_SUCCESS;
}
