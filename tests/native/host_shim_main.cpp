// Drives the C++ host classes (hero_amd/host/hero_host.h) the way the reference's main.cpp drives
// its own: used by tests/test_gpu_host_shims.py.
//   host_shim_main correct <reads.fasta> <k> <m%> <i%> <l> <g|a> <out.fasta> [subject_chunk]
//   host_shim_main aln     <reads.fasta> <k> <m%> <i%> <l> <g|a> <out.aln>   [subject_chunk]
//   host_shim_main band    <g|a> <cases.txt> <out.txt>   (lines "S Q band flag" -> "band cigar")
// subject_chunk > 0 streams the subjects through SubjectDataset chunks instead of self-alignment.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "../../hero_amd/host/hero_host.h"

using namespace hero_host;

// correct: -c, corrected FASTA; aln: the .aln text; mapping: -c --mapping, the coverage histogram
// file (main.cpp:82-87 writes it instead of calling consensus)
static int run_path(bool correct, bool mapping, int argc, char **argv) {
	if (argc < 9) return 2;
	ReadBatch reads;
	if (!ReadBatch::loadFasta(argv[2], reads)) return 3;
	Config cfg;
	cfg.setKmer(std::atoi(argv[3]));
	cfg.setMaxMismatchRatePercent(std::atof(argv[4]));
	cfg.setMaxIndelRatePercent(std::atof(argv[5]));
	cfg.setMinimumOverlap(std::atoll(argv[6]));
	cfg.setAffine(argv[7][0] == 'a');
	cfg.setCallConsensus(correct);
	cfg.setMapping(mapping);
	const int64_t chunk = argc > 9 ? std::atoll(argv[9]) : 0;
	Device dev(cfg);
	QueryDataset q(dev, reads);
	HashTableLongKey table(&q);
	table.createHashTables();
	table.insertQueryDataset(&q);
	SubjectReadAligner aligner(&q, &table);
	if (chunk > 0) {
		SubjectDataset subjects(reads, chunk);
		aligner.start(subjects);
	} else {
		aligner.startSelf();
	}
	if (mapping) {
		std::ofstream(argv[8], std::ios_base::binary | std::ios_base::trunc).close();
		q.getCoverageAndWriteResults(argv[8]);
		return 0;
	}
	std::ofstream out(argv[8], std::ios_base::binary | std::ios_base::trunc);
	if (correct) {
		q.callConsensus();
		out << q.queryOutText();
	} else {
		out << aligner.alignments();
	}
	return 0;
}

static int run_band(int argc, char **argv) {
	if (argc < 5) return 2;
	Config cfg;
	cfg.setAffine(argv[2][0] == 'a');
	Device dev(cfg);
	BandAligner ba(dev);
	std::ifstream in(argv[3]);
	std::ofstream out(argv[4], std::ios_base::binary | std::ios_base::trunc);
	std::string line;
	while (std::getline(in, line)) {
		std::istringstream ls(line);
		std::string S, Q;
		int64_t band;
		int flag;
		if (!(ls >> S >> Q >> band >> flag)) continue;
		if (S == "-") S.clear();
		if (Q == "-") Q.clear();
		ba.setBand(band);
		ba.setSequences(S, 0, (int64_t) S.size(), Q, 0, (int64_t) Q.size());
		if (flag == 1) ba.setAlignLeft();
		if (flag == 2) ba.setAlignRight();
		ba.calculateAlignment();
		out << ba.m_band << ' ';
		for (size_t i = 0; i < ba.getCigarCount()->size(); i++) out << (*ba.getCigarCount())[i] << (*ba.getCigarType())[i];
		out << '\n';
	}
	return 0;
}

int main(int argc, char **argv) {
	if (argc < 2) return 2;
	const std::string mode = argv[1];
	try {
		if (mode == "correct") return run_path(true, false, argc, argv);
		if (mode == "aln") return run_path(false, false, argc, argv);
		if (mode == "mapping") return run_path(true, true, argc, argv);
		if (mode == "band") return run_band(argc, argv);
	} catch (const std::exception &e) {
		std::fprintf(stderr, "%s\n", e.what());
		return 1;
	}
	return 2;
}
