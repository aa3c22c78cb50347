// Driver around the reference front end (pareas-ast, compiled from its own
// sources under the reference checkout by oracle/Makefile into oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY: it writes the DepthTree arrays the reference front
// end produces for a source file, so the tests can pin the shipped front end
// (ptilopsis_amd/csrc/frontend.cpp) against it.  Output layout (little endian):
//   u64 n_nodes, u64 max_depth, then u8 types[n], u8 rtypes[n], i32 parents[n],
//   i32 depth[n], i32 child_idx[n], u32 node_data[n].
#include <cstdio>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <memory>

#include "codegen/lexer.hpp"
#include "codegen/parser.hpp"
#include "codegen/astnode.hpp"
#include "codegen/exception.hpp"
#include "codegen/depthtree.hpp"
#include "codegen/symtab.hpp"

int main(int argc, char **argv) {
    if (argc != 3) { std::cerr << "usage: pareas_dump <in> <out>\n"; return 2; }
    try {
        std::ifstream in(argv[1]);
        if (!in) { std::cerr << "cannot open " << argv[1] << "\n"; return 2; }
        Lexer lexer(in);
        SymbolTable symtab;
        Parser parser(lexer, symtab);
        std::unique_ptr<ASTNode> node(parser.parse());
        node->resolveType();
        DepthTree tree(node.get());
        FILE *f = std::fopen(argv[2], "wb");
        uint64_t n = tree.filledNodes(), md = tree.maxDepth();
        std::fwrite(&n, 8, 1, f); std::fwrite(&md, 8, 1, f);
        std::fwrite(tree.getNodeTypes().data(), 1, n, f);
        std::fwrite(tree.getResultingTypes().data(), 1, n, f);
        std::fwrite(tree.getParents().data(), 4, n, f);
        std::fwrite(tree.getDepth().data(), 4, n, f);
        std::fwrite(tree.getChildren().data(), 4, n, f);
        std::fwrite(tree.getNodeData().data(), 4, n, f);
        std::fclose(f);
    } catch (const ParseException &e) {
        std::cerr << "Parse error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
