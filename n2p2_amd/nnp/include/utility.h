// String and file helpers of libnnp that callers use (utility.h of the
// reference): printf-style formatting, splitting, column-file headers.
#ifndef NNP_UTILITY_H
#define NNP_UTILITY_H

#include <cstddef>
#include <fstream>
#include <string>
#include <vector>

namespace nnp
{

std::vector<std::string> split(std::string const& input, char delimiter = ' ');
std::string trim(std::string const& line, std::string const& whitespace = " \t");
std::string reduce(std::string const& line,
                   std::string const& whitespace = " \t",
                   std::string const& fill = " ");
std::string pad(std::string const& input, std::size_t num, char fill = ' ',
                bool right = true);
std::string strpr(const char* format, ...);
std::string cap(std::string word);
// Column-file header: title block, column table and the numbered column
// names, the layout of the reference's output files (utility.cpp).
std::vector<std::string> createFileHeader(
    std::vector<std::string> const& title,
    std::vector<std::size_t> const& colSize,
    std::vector<std::string> const& colName,
    std::vector<std::string> const& colInfo,
    char const& commentChar = '#');
void appendLinesToFile(std::ofstream& file,
                       std::vector<std::string> const lines);
void appendLinesToFile(FILE* const& file,
                       std::vector<std::string> const lines);
double pow_int(double x, int n);

}  // namespace nnp

#endif
