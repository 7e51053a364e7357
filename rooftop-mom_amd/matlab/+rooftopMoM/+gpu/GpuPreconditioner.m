classdef GpuPreconditioner < rooftopMoM.Product
% GpuPreconditioner: A preconditioner of a Problem applied on the GPU, a drop-in for rooftopMoM.Preconditioner
%   (src/+rooftopMoM/Preconditioner.m:7-70): the reference's circulant preconditioners ('kernel', 'builder',
%   'bandwidth', 'modifier', as Preconditioner.m:29-34), or 'kind' 'nearField' (3x3-cell neighbourhoods) or 'jacobi'.
%       preconditioner = rooftopMoM.gpu.GpuPreconditioner(problem, 'kind', 'nearField', 'precision', 'mixed');
%       preconditioner = rooftopMoM.gpu.GpuPreconditioner(problem, 'kernel', 'rChan', 'builder', 'blockwise');

    properties (Access = private)
        handle_ uint64;
    end

    methods
        function obj = GpuPreconditioner(problem, options)
            arguments
                problem (1, 1) rooftopMoM.Problem
                options.kind (1, :) char {mustBeMember(options.kind, {'circulant', 'nearField', 'jacobi'})} = ...
                    'circulant'
                options.kernel (1, :) char {mustBeMember(options.kernel, ...
                    {'strang', 'tChan', 'rChan', 'huckle', 'superoptimal'})} = 'rChan'
                options.builder (1, :) char {mustBeMember(options.builder, {'partwise', 'blockwise', 'hybrid'})} = ...
                    'partwise'
                options.bandwidth double = []
                options.modifier function_handle = function_handle.empty
                options.precision (1, :) char {mustBeMember(options.precision, {'c128', 'c64', 'mixed'})} = 'c128'
            end

            precision = find(strcmp(options.precision, {'c128', 'c64', 'mixed'})) - 1;
            problemHandle = rooftopMoM.gpu.problemHandle(problem);
            cleanup = onCleanup(@() rooftop_mex('free', 'problem', problemHandle));

            switch options.kind
                case 'nearField'
                    obj.handle_ = rooftop_mex('precondition', problemHandle, 2, precision);
                case 'jacobi'
                    obj.handle_ = rooftop_mex('precondition', problemHandle, 1, precision);
                otherwise
                    kernel = find(strcmp(options.kernel, {'strang', 'tChan', 'rChan', 'huckle', 'superoptimal'})) - 1;
                    builder = find(strcmp(options.builder, {'partwise', 'blockwise', 'hybrid'})) - 1;
                    bandwidth = [0, 0];

                    if ~isempty(options.bandwidth)
                        if numel(options.bandwidth) ~= 2 || any(options.bandwidth <= 0) || ...
                                any(options.bandwidth ~= round(options.bandwidth))
                            error('rooftopMoM:preconditioner:invalidBandwidth', ...
                                'The bandwidth is a pair of positive integers.');
                        end

                        bandwidth = double(options.bandwidth);
                    end

                    modifier = 0;

                    if ~isempty(options.modifier)
                        modifier = find(strcmp(func2str(options.modifier), {'real', 'imag'}));
                    end

                    obj.handle_ = rooftop_mex('circulant', problemHandle, kernel, builder, bandwidth, modifier, ...
                        precision);
            end

            obj.problem_ = problem;
            obj.applicator_ = rooftopMoM.gpu.GpuApplicator(obj.handle_, min(precision, 1));
        end

        function delete(obj)
            if ~isempty(obj.handle_)
                rooftop_mex('free', 'product', obj.handle_);
            end
        end

        function handle = handle(obj)
            handle = obj.handle_;
        end
    end
end
